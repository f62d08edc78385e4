/*
 * grmp.h -- C ABI of libgrmp_hip.so, the MI355X (gfx950) assembly-and-solve
 * engine behind GradientRobustMultiPhysics.jl's hot path.
 *
 * The entry points replace, one for one, the reference functions named in
 * each comment (paths relative to the reference root).  They take plain
 * pointers and sizes only.  Index conventions follow Julia, the reference's
 * host language: index arrays are 1-based and matrices are column-major, so
 * a Julia caller passes xgrid[Coordinates] ([dim, nnodes]), xgrid[CellNodes]
 * ([nloc, ncells]) and FES[CellDofs] ([ndofs_loc, ncells]) unchanged.
 *
 * Ownership: host pointers are borrowed for the duration of a call; the
 * library copies what it keeps.  Device buffers belong to the opaque handles
 * and are released by the matching *_destroy.  Calls on one context are
 * serialized by the caller; different contexts are independent.
 *
 * Every function returns GRMP_OK (0) or an error code; grmp_last_error()
 * gives the message of the last failure on that context.
 */
#ifndef GRMP_H
#define GRMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define GRMP_OK 0
#define GRMP_ERR_ARG 1          /* malformed argument (cf. @assert, bilinearform.jl:39-40) */
#define GRMP_ERR_OOM 2          /* device allocation failed */
#define GRMP_ERR_HIP 3          /* HIP runtime error */
#define GRMP_ERR_NOCONV 4       /* max iterations reached (cf. solvers.jl:1398-1400) */
#define GRMP_ERR_NONFINITE 5    /* NaN/Inf met in a Krylov loop */
#define GRMP_ERR_ADMISSION 6    /* request larger than the server's memory budget */
#define GRMP_ERR_UNSUPPORTED 7  /* element / operator combination not built */
#define GRMP_ERR_BUSY 8         /* result not ready yet (server poll) */

/* --------------------------------------------------------- enumerations */
/* finite element types (files under src/fedefs/) */
#define GRMP_FE_H1P1 1          /* H1P1{ncomp}      src/fedefs/h1_p1.jl */
#define GRMP_FE_H1P2 2          /* H1P2{ncomp,2}    src/fedefs/h1_p2.jl */
#define GRMP_FE_H1BR 3          /* H1BR{2}          src/fedefs/h1v_br.jl */
#define GRMP_FE_L2P0 4          /* L2P0{1}          src/fedefs/l2_p0.jl */

/* bilinear forms (src/pdeoperators.jl) */
#define GRMP_BLF_LAPLACE 1      /* (kappa grad u, grad v): LaplaceOperator :154-169, or
                                   BilinearForm([Gradient,Gradient], kappa-action;
                                   APT_SymmetricBilinearForm) :355-409 */
#define GRMP_BLF_HOOKE2D 2      /* (C eps(u), eps(v)): HookStiffnessOperator2D :256-276 */
#define GRMP_BLF_MASS 3         /* (kappa u, v): ReactionOperator :176-203 */
#define GRMP_BLF_DIVP0 4        /* LagrangeMultiplier(Divergence) :214-226 (with
                                   transposed copy) */
#define GRMP_BLF_NLDIFF 5       /* NonlinearForm(Gradient, [Identity, Gradient], [1,1], kernel, [2,3])
                                   (:590-635) with kernel(u, grad u) = alpha(u) grad u,
                                   alpha(u) = params[0] + params[1] u + params[2] u^2: the Newton matrix
                                   of full_assemble! (src/assemblypatterns/nonlinearform.jl:44-246, rows
                                   = test functions as transposed_assembly sets it) at the iterate
                                   `current`; its pattern holds every local (test, ansatz) pair */

/* how the coefficient enters (src/assemblypatterns/bilinearform.jl:306-313) */
#define GRMP_COEF_FACTOR 0      /* NoAction: constant coefficient multiplies itemfactor */
#define GRMP_COEF_CELL 1        /* item-dependent Action: result = kappa[cell] .* input */

/* linear forms */
#define GRMP_LF_CONST 1         /* LinearForm(Identity, DataFunction(f)) :694-705 */
#define GRMP_LF_RECO_RT0 2      /* LinearForm(ReconstructionIdentity{HDIVRT0{2}}, f) on H1BR{2},
                                   f = grad(phi) + curl(psi) (x-dependent, see grmp_lform) */
#define GRMP_LF_POLY 3          /* LinearForm(Identity, f) with f a polynomial of degree <= 4 per
                                   component (coef, ncoef) on P1/P2 triangles */
#define GRMP_LF_NEWTON 4        /* the right-hand side full_assemble! adds for GRMP_BLF_NLDIFF at
                                   `current` (nonlinearform.jl:190-200, 218-224):
                                   (J [u, grad u] - kernel(u, grad u)) . grad v, alpha from f[0..2] */

/* value precision flags of an assembled matrix */
#define GRMP_VALUES_FP64 1
#define GRMP_VALUES_FP32 2

/* solver choices (replaces the LinearSolve/UMFPACK call, src/solvers.jl:656-660) */
#define GRMP_SOLVE_CG 1
#define GRMP_SOLVE_MINRES 2
#define GRMP_SOLVE_BICGSTAB 3   /* nonsymmetric systems (Newton matrices), FP64, Jacobi */
#define GRMP_PREC_JACOBI 1
#define GRMP_PRECISION_FP64 1   /* plain FP64 Krylov */
#define GRMP_PRECISION_FP32_IR 2/* FP32 matrix/SpMV/preconditioner inside FP64 refinement */

/* ---------------------------------------------------------------- types */
typedef struct grmp_ctx grmp_ctx;
typedef struct grmp_mesh grmp_mesh;
typedef struct grmp_space grmp_space;
typedef struct grmp_vector grmp_vector;
typedef struct grmp_matrix grmp_matrix;
typedef struct grmp_server grmp_server;
typedef struct grmp_mg grmp_mg;

typedef struct grmp_form {
    int kind;                 /* GRMP_BLF_* */
    int coef_mode;            /* GRMP_COEF_* */
    int quadorder;            /* as prepare_assembly! chooses it (assemblypatterns.jl:559-565) */
    int row_block, col_block; /* block of FEMatrix (0-based) */
    double factor;            /* PDEOperator factor (NoAction coefficient) */
    double params[4];         /* Hooke: lambda, mu */
    const grmp_vector *cell_coef; /* per-cell kappa for GRMP_COEF_CELL, else NULL */
    const grmp_vector *current;   /* GRMP_BLF_NLDIFF: the iterate (read at every assemble) */
    int64_t current_offset;       /* its block offset */
} grmp_form;

typedef struct grmp_lform {
    int kind;                 /* GRMP_LF_CONST / GRMP_LF_RECO_RT0 */
    int quadorder;            /* RECO_RT0: 2 + bonus order of f (Stroud rule above 2) */
    double factor;
    double f[4];              /* constant data, ncomp values */
    int nmodes;               /* RECO_RT0: phi and psi are sums of nmodes sine modes each */
    int pad_;
    const double *modes;      /* [2*nmodes][4] = (amplitude, p, q, phase): A sin(2 pi (p x + q y) + t),
                                 phi's modes first; host pointer, copied */
    double poly[4];           /* phi += poly0 x^3 + poly1 x y^2 + poly2 y^2 */
    const grmp_vector *current;   /* GRMP_LF_NEWTON: the iterate */
    int64_t current_offset;
    const double *coef;       /* GRMP_LF_POLY: [ncomp][ncoef] over 1, x, y, x^2, xy, y^2, x^3, x^2 y,
                                 x y^2, y^3, x^4, x^3 y, x^2 y^2, x y^3, y^4; host pointer, copied */
    int ncoef;                /* <= 15 */
    int pad2_;
} grmp_lform;

typedef struct grmp_solve_opts {
    int method;               /* GRMP_SOLVE_CG (symmetric positive definite systems) /
                                 GRMP_SOLVE_MINRES (saddle points: free rows without a diagonal
                                 are constraint rows, negated to make the system symmetric) /
                                 GRMP_SOLVE_BICGSTAB (nonsymmetric, FP64 only) */
    int precond;              /* GRMP_PREC_JACOBI (MINRES: 1/|A_ii|, and on constraint rows
                                 1 / sum_j A_ij^2 / A_jj) */
    int precision;            /* GRMP_PRECISION_FP64 / GRMP_PRECISION_FP32_IR */
    double rtol;              /* CG: ||b - Ax|| <= rtol ||b - A x0|| over free dofs;
                                 MINRES: the same in the preconditioner's norm */
    int64_t maxit;            /* total Krylov iterations */
    double inner_rtol;        /* FP32_IR CG: reliable-update threshold, the FP64 residual is
                                 recomputed when the FP32 one has dropped by this factor;
                                 FP32_IR MINRES: relative tolerance of each FP32 inner solve */
    int64_t inner_maxit;      /* FP32_IR MINRES: iteration cap of each inner solve */
    int check_every;          /* MINRES: status poll period in iterations (results do not
                                 depend on it) */
    int initial_guess;        /* BiCGStab: 1 = start from the free entries of x (the previous
                                 Newton iterate); the target stays rtol times the residual of
                                 the zero start.  CG and MINRES start from zero and reject 1 */
} grmp_solve_opts;

typedef struct grmp_solve_info {
    int status;
    int outer;                /* FP64 residual evaluations (refinement refreshes) */
    int64_t iterations;       /* Krylov iterations (inner ones in FP32_IR) */
    double relres;            /* final FP64 relative residual */
    double seconds;
} grmp_solve_info;

/* ------------------------------------------------------------- context */
int grmp_ctx_create(int device, grmp_ctx **out);
int grmp_ctx_destroy(grmp_ctx *ctx);
const char *grmp_last_error(const grmp_ctx *ctx);
int grmp_ctx_synchronize(grmp_ctx *ctx);
/* Pattern builds of more than max_nnz nonzeros fail with GRMP_ERR_UNSUPPORTED
 * (the total is summed in 64 bits before the int32 row-pointer scan); the
 * limit is at most and by default 2^31 - 1, the int32 CSR bound.  A memory
 * guard with no counterpart in the reference (Int64 CSC). */
int grmp_ctx_set_max_nnz(grmp_ctx *ctx, int64_t max_nnz);
/* library build/ABI version: major*10000 + minor*100 + patch */
int grmp_version(void);

/* ---------------------------------------------------- device vectors */
int grmp_vector_create(grmp_ctx *ctx, int64_t n, grmp_vector **out);
int grmp_vector_destroy(grmp_vector *v);
int64_t grmp_vector_length(const grmp_vector *v);
int grmp_vector_upload(grmp_vector *v, const double *host);
int grmp_vector_download(const grmp_vector *v, double *host);
int grmp_vector_fill(grmp_vector *v, double value);
/* page-locked host buffers for results (the solution arrays a Julia caller
 * wraps with unsafe_wrap): device-to-host copies into them need no staging.
 * No counterpart in the reference, which keeps every array on the host. */
int grmp_host_alloc(grmp_ctx *ctx, size_t bytes, void **out);
int grmp_host_free(void *p);
/* v[idx[k]] = values[k] (idx 1-based); the b[fixed] = penalty*target write of
 * solve_direct! (src/solvers.jl:649-650) */
int grmp_vector_scatter(grmp_vector *v, int64_t count, const int64_t *idx, const double *values);
/* seeded log-normal field exp(mu + sigma Z), Z ~ N(0,1) from Philox-4x32-10
 * keyed by seed and counted by entry index (Box-Muller) */
int grmp_vector_lognormal(grmp_vector *v, uint64_t seed, double mu, double sigma);
/* y += a x (lifting of Dirichlet values around a solve with homogeneous
 * data; no counterpart of its own in the reference) */
int grmp_vector_axpy(double a, const grmp_vector *x, grmp_vector *y);

/* --------------------------------------------------------------- meshes */
/* ExtendableGrid upload: Coordinates [dim,nnodes], CellNodes [nloc,ncells]
 * (1-based).  dim 2 (Triangle2D, nloc 3) or 3 (Tetrahedron3D, nloc 4). */
int grmp_mesh_upload(grmp_ctx *ctx, int dim, int64_t nnodes, const double *coords,
                     int64_t ncells, int nloc, const int32_t *cellnodes, grmp_mesh **out);
/* Unit square with N x N squares, each split along its (0,0)-(1,1) diagonal,
 * generated on the device (node id j(N+1)+i, coordinates i/N). */
int grmp_mesh_unitsquare(grmp_ctx *ctx, int N, grmp_mesh **out);
/* Unit cube with n^3 cubes in 6 tetrahedra each (Kuhn split), interior
 * nodes jittered by +-jitter*h per coordinate, boundary nodes tangentially. */
int grmp_mesh_unitcube(grmp_ctx *ctx, int n, double jitter, uint64_t seed, grmp_mesh **out);
int grmp_mesh_destroy(grmp_mesh *m);
int grmp_mesh_info(const grmp_mesh *m, int *dim, int64_t *nnodes, int64_t *ncells);
int grmp_mesh_download(const grmp_mesh *m, double *coords, int32_t *cellnodes);
/* per-cell volumes (ExtendableGrids CellVolumes) */
int grmp_mesh_cellvolumes(const grmp_mesh *m, double *out);
/* Face adjacency of a triangle mesh, built on the device when the mesh is
 * uploaded: ExtendableGrids' CellFaces [3,ncells] (1-based, column-major, so
 * cellfaces[3c+f]), CellFaceSigns [3,ncells] and a 0/1 boundary flag per face
 * (the faces BFaceFaces lists).  Local face f joins local nodes f and f%3+1;
 * faces are numbered by first appearance in cell order and oriented by that
 * appearance.  Replaces instantiate(xgrid, CellFaces) / CellFaceSigns /
 * BFaceFaces (ExtendableGrids; call sites src/dofmaps.jl:236,401-460,
 * src/fedefs/h1v_br.jl:151-161).  Any output may be NULL; *nfaces is always
 * written. */
int grmp_mesh_faces(const grmp_mesh *m, int64_t *nfaces, int32_t *cellfaces, int8_t *cellfacesigns,
                    uint8_t *bface);

/* ------------------------------------------------------------- FESpace */
/* FESpace from the caller's dofmap: FES[CellDofs] ([ndofs_loc, ncells],
 * 1-based; src/dofmaps.jl:201-363).  Replaces FESpace{FEType}(xgrid)
 * (src/finiteelements.jl:72-112) on the device side. */
int grmp_space_upload(grmp_ctx *ctx, const grmp_mesh *mesh, int fe, int ncomp, int64_t ndofs,
                      int ndofs_loc, const int32_t *celldofs, grmp_space **out);
/* FESpace with the dofmap generated on the device: P1 ("N1") on any mesh;
 * P2 ("N1F1") on a grmp_mesh_unitsquare mesh, faces numbered by first
 * appearance in cell order with local faces [1 2; 2 3; 3 1]. */
int grmp_space_create(grmp_ctx *ctx, const grmp_mesh *mesh, int fe, int ncomp, grmp_space **out);
int grmp_space_destroy(grmp_space *s);
int grmp_space_info(const grmp_space *s, int64_t *ndofs, int *ndofs_loc);
int grmp_space_download_celldofs(const grmp_space *s, int32_t *celldofs);
/* BFaceDofs of all boundary faces of a unit-square/unit-cube mesh (the dofs
 * HomogeneousDirichletBoundary fixes, src/boundarydata.jl:198-257); 1-based,
 * sorted.  Call with out == NULL to get the count. */
int grmp_space_boundary_dofs(const grmp_space *s, int64_t *count, int64_t *out);

/* ------------------------------------------------------- system matrix */
/* FEMatrix over nblocks FESpaces (src/fematrix.jl:186-213). */
int grmp_matrix_create(grmp_ctx *ctx, int nblocks, const grmp_space *const *spaces,
                       grmp_matrix **out);
int grmp_matrix_destroy(grmp_matrix *A);
/* register a BilinearForm for block (row_block, col_block); forms are
 * assembled in registration order, as operators of one block are. */
int grmp_matrix_add_form(grmp_matrix *A, const grmp_form *form);
/* fixed dofs (1-based, global) and penalty: A[d,d] = penalty after assembly
 * (apply_penalties!, src/fematrix.jl:349-355; src/solvers.jl:647-652) */
int grmp_matrix_set_fixed(grmp_matrix *A, int64_t nfixed, const int64_t *fixed, double penalty);
/* Builds the sparsity pattern the reference's first assembly creates: the
 * set of (i,j) that received a nonzero local contribution (_addnz,
 * src/fematrix.jl:54-65) plus penalized diagonals, with the element-to-slot
 * map.  Uses the current coefficient values. */
int grmp_matrix_build_pattern(grmp_matrix *A);
/* assemble!(A::FEMatrixBlock, AP) for all registered forms
 * (src/assemblypatterns/bilinearform.jl:92-400).  values: GRMP_VALUES_* mask. */
int grmp_matrix_assemble(grmp_matrix *A, int values);
int grmp_matrix_info(const grmp_matrix *A, int64_t *n, int64_t *nnz, int *max_row);
/* A.entries.cscmatrix: colptr[n+1], rowval[nnz] (1-based), nzval[nnz] */
int grmp_matrix_export_csc(const grmp_matrix *A, int64_t *colptr, int64_t *rowval, double *nzval);
/* y = A x with the FP64 or FP32 values (x, y FP64 device vectors) */
int grmp_matrix_spmv(grmp_matrix *A, int values, const grmp_vector *x, grmp_vector *y);
/* Batched assembly of many meshes: the FP64 values of `count` single-form
 * matrices (any meshes, spaces and form kinds) in one launch of a persistent
 * kernel over a queue of (matrix, row block) tiles, bitwise equal to
 * assembling each (bilinearform.jl:226-377 per item).  Builds missing
 * patterns first.  No counterpart in the reference, which assembles one
 * operator at a time. */
int grmp_matrix_assemble_batch(int count, grmp_matrix *const *mats);

/* seconds of the last assemble call's kernels (HIP events) */
double grmp_matrix_last_assembly_seconds(const grmp_matrix *A);

/* Launch shape of the element assembly (no counterpart in the reference,
 * whose element loop is serial): variant 1 walks each row's cells in one
 * thread, 2 stages the row block's cell work over all threads of a
 * workgroup, 3 is 2 with several cells' gathers in flight per thread (form
 * kinds whose rows need only the cell's corners and coefficient; others run
 * 2); rows_per_block the rows one workgroup owns; xcd_remap 1 gives
 * each XCD a contiguous share of the row blocks.  0 (-1 for xcd_remap)
 * keeps the built-in choice.  Values are bitwise the same for every shape.
 * grmp_matrix_set_tune overrides the context's table for one matrix (NULL
 * resets); grmp_matrix_tune_key returns the table key of a single-form
 * matrix: form kind, quadrature points and nnz/row bucket (<= 8, 16, 32, 64,
 * above).  grmp_tune_load reads a table written by the autotuner (JSON; path
 * NULL: tuning_gfx950.json beside the library or $GRMP_TUNING), which
 * grmp_ctx_create also loads when present; grmp_tune_count is its size. */
typedef struct grmp_asm_tune {
    int32_t variant;
    int32_t rows_per_block;
    int32_t xcd_remap;
} grmp_asm_tune;
int grmp_matrix_set_tune(grmp_matrix *A, const grmp_asm_tune *t);
int grmp_matrix_tune_key(const grmp_matrix *A, int *kind, int *nq, int *bucket);
int grmp_tune_load(grmp_ctx *ctx, const char *path);
int grmp_tune_count(const grmp_ctx *ctx);

/* The right-hand side of one BackwardEuler step of advance!
 * (src/solvers_timedependent.jl:339-359 with the penalty of :452-461):
 * b = f + inv_dt * M u on the free rows of the system matrix S, and
 * b_d = penalty * u_d on its fixed rows.  M: the time-derivative mass matrix
 * (FP64 values); f, u, b of S's length; b may be f but not u. */
int grmp_timestep_rhs(const grmp_matrix *S, const grmp_matrix *M, const grmp_vector *f, const grmp_vector *u,
                      double inv_dt, grmp_vector *b);

/* ---------------------------------------------------------- LinearForm */
/* assemble!(b::FEVectorBlock, AP) for a LinearForm, added into b at offset
 * (src/assemblypatterns/linearform.jl:47-251) */
int grmp_assemble_linearform(grmp_ctx *ctx, const grmp_space *space, const grmp_lform *form,
                             grmp_vector *b, int64_t offset);

/* ------------------------------------------------------ ItemIntegrator */
/* evaluate(ItemIntegrator([Identity]), u): per-cell integrals of the FE
 * function u (coefficients at offset), out [ncomp, ncells] column-major
 * (src/assemblypatterns/itemintegrator.jl:169-360).  H1P1, H1P2, L2P0. */
int grmp_integrate_identity(grmp_ctx *ctx, const grmp_space *space, int quadorder,
                            const grmp_vector *u, int64_t offset, double *out);

/* evaluate(L2ErrorIntegrator(g, Identity; quadorder, factor), u) and, with
 * ncoef = 0, evaluate(L2NormIntegrator(ncomp, Identity), u)
 * (src/assemblypatterns/itemintegrator.jl:40-95, evaluate! :169-307,
 * evaluate :323-360).  Per cell sum_q w_q |T| sum_k (g_k(x_q) - factor
 * u_k(x_q))^2 with g_k the polynomial coef[k*ncoef .. k*ncoef+ncoef) over the
 * monomials 1, x, y, x^2, xy, y^2, x^3, x^2 y, x y^2, y^3.  cellvals
 * ([ncells], optional) gets evaluate!'s per-item values, *total (optional)
 * their sum, the squared L2 error.  H1P1/H1P2 (1-2 components), H1BR{2},
 * L2P0 on triangles; quadorder as the reference's triangle rules (<= 2
 * closed form, else Stroud). */
int grmp_integrate_l2error(grmp_ctx *ctx, const grmp_space *space, const grmp_vector *u, int64_t offset,
                           const double *coef, int ncoef, int quadorder, double factor, double *cellvals,
                           double *total);

/* The same, allocating nothing: the per-cell values and partials live in a
 * file-scope device array of the library (one per GPU, held until the
 * process ends), so callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc) can
 * evaluate errors.  No counterpart in the reference.  At most 4,194,304
 * cells (GRMP_ERR_ARG above); calls that overlap on one GPU share the array
 * and race. */
int grmp_integrate_l2error_noalloc(grmp_ctx *ctx, const grmp_space *space, const grmp_vector *u, int64_t offset,
                                   const double *coef, int ncoef, int quadorder, double factor, double *cellvals,
                                   double *total);

/* ------------------------------------------------------ Dirichlet data */
/* Boundary values of an H1P1 / H1P2 space with 1 or 2 components (triangle
 * meshes) for the polynomial data g_c = sum_k coef[c ncoef + k] m_k,
 * m = (1, x, y, x^2, xy, y^2, x^3, x^2 y, x y^2, y^3), ncoef <= 10, coef
 * holding ncomp * ncoef values; written to target[offset + dof] for the dofs
 * of all boundary faces (other entries untouched).
 * InterpolateDirichletBoundary (src/boundarydata.jl:198-257): g at the dof
 * points. */
int grmp_boundary_interpolate(const grmp_space *space, const double *coef, int ncoef, grmp_vector *target,
                              int64_t offset);
/* BestapproxDirichletBoundary (src/boundarydata.jl:259-352): the L2
 * projection of g onto the traces on the boundary faces (face mass matrix and
 * load of g with a Gauss rule of order quadorder <= 7), solved to relative
 * residual rtol; iterations / relres may be NULL. */
int grmp_boundary_bestapprox(const grmp_space *space, const double *coef, int ncoef, int quadorder, double rtol,
                             grmp_vector *target, int64_t offset, int64_t *iterations, double *relres);
/* b_d = penalty * target_d on the fixed rows of A (solve_direct!,
 * src/solvers.jl:647-652, with Target from boundarydata!). */
int grmp_matrix_penalize_rhs(const grmp_matrix *A, const grmp_vector *target, grmp_vector *b);

/* nodevalues!(target, source, FE, Identity) as writeCSV! uses it
 * (src/dataexport.jl:22-77, src/interpolations.jl:704-812,
 * src/fedefs/l2_p0.jl:52-68): per node the values at that vertex of the cells
 * around it, summed in cell order and divided by their number.  H1P1, H1P2,
 * H1BR, L2P0 on triangles; out [ncomp, nnodes] column-major (host). */
int grmp_nodevalues(const grmp_space *space, const grmp_vector *u, int64_t offset, double *out);

/* ---------------------------------------------------------------- solve */
void grmp_solve_opts_default(grmp_solve_opts *o);
/* Solves the penalized system A x = b (b[fixed] = penalty*value) as
 * solve_direct! does (src/solvers.jl:622-686), with fixed dofs eliminated. */
int grmp_solve(grmp_matrix *A, const grmp_vector *b, grmp_vector *x, const grmp_solve_opts *opts,
               grmp_solve_info *info);

/* Multigrid-preconditioned FP64 CG (V-cycle with Chebyshev smoothing, the
 * served solver) on a caller-built hierarchy of generated unit squares:
 * levels[0] is the system over spaces[0] (H1P1 or H1P1{2} on N x N, or H1P2
 * followed by H1P1 on the same N), each next level a matrix with the same
 * forms, fixed boundary dofs and pattern on ceil(N/2) x ceil(N/2), down to
 * at most 64 dofs.  The job assembles every level itself: forms with
 * GRMP_COEF_CELL take kappa0 on level 0 and its geometric means on the
 * coarser cells.  Homogeneous Dirichlet data (x = 0 on fixed rows).  No
 * counterpart in the reference, which factorizes (src/solvers.jl:55,
 * 656-660); used for the systems of TimeControlSolver. */
int grmp_mg_solve(grmp_ctx *ctx, int nlev, const grmp_matrix *const *levels, const grmp_space *const *spaces,
                  const grmp_vector *kappa0, const grmp_vector *b, grmp_vector *x, double rtol, int64_t maxit,
                  grmp_solve_info *info);
/* The same in parts, for many right-hand sides with one operator (the reuse
 * of the factorization in advance!, src/solvers_timedependent.jl:337,
 * :463-466): create assembles the hierarchy (levels, spaces and kappa0 must
 * outlive the handle), apply solves for one b, destroy frees the handle. */
int grmp_mg_create(grmp_ctx *ctx, int nlev, const grmp_matrix *const *levels, const grmp_space *const *spaces,
                   const grmp_vector *kappa0, grmp_mg **out);
int grmp_mg_apply(grmp_mg *m, const grmp_vector *b, grmp_vector *x, double rtol, int64_t maxit,
                  grmp_solve_info *info);
int grmp_mg_destroy(grmp_mg *m);

/* The Stokes system of grmp_solve's MINRES (src/solvers.jl:622-686 replaced
 * for IncompressibleNavierStokesProblem(2; nonlinear=false),
 * src/pdeprototypes.jl:17-55) with a block-diagonal preconditioner: one
 * V-cycle of `velocity` -- a hierarchy (grmp_mg_create) of the viscosity
 * times the scalar P1 Laplacian on the velocity's mesh, with its boundary
 * dofs fixed -- for the node rows of each of the ncomp velocity components
 * (rows [k n, (k+1) n), n the hierarchy's level-0 size: the P1 part of
 * H1BR{2}, on which the viscous block is exactly that Laplacian), 1/|A_ii| on
 * the other velocity rows (face bubbles) and the Schur-complement diagonal on
 * the constraint rows.  FP64, zero start.  No counterpart in the reference,
 * which factorizes. */
int grmp_solve_minres_mg(grmp_matrix *A, const grmp_vector *b, grmp_vector *x, const grmp_solve_opts *opts,
                         grmp_mg *velocity, int ncomp, grmp_solve_info *info);

/* Algebraic multigrid as the preconditioner of FP64 CG, for symmetric
 * positive definite systems on any mesh (uploaded meshes, tetrahedra: the
 * multigrid above needs generated unit squares).  The hierarchy is built on
 * the device from the assembled FP64 matrix: two passes of pairwise matching
 * per level (aggregates of up to four rows), Galerkin coarse operators,
 * Chebyshev smoothing with FP32-stored values, the dense inverse on the
 * coarsest level.  Fixed (penalized) dofs are eliminated (x_d = b_d / A_dd),
 * so b is the penalized right-hand side of grmp_solve.  The handle reads A's
 * pattern and values: keep A alive, and for grmp_amg_apply unchanged (create
 * again after it is reassembled; grmp_solve_bicgstab_amg below also takes a
 * hierarchy of an earlier assembly of A's pattern).  No counterpart in the
 * reference, which factorizes (src/solvers.jl:55, 656-660). */
typedef struct grmp_amg grmp_amg;
int grmp_amg_create(grmp_matrix *A, grmp_amg **out);
int grmp_amg_apply(grmp_amg *m, const grmp_vector *b, grmp_vector *x, double rtol, int64_t maxit,
                   grmp_solve_info *info);
/* levels, rows of the coarsest level, operator complexity (summed nonzeros / level-0 nonzeros) */
int grmp_amg_info(const grmp_amg *m, int *nlev, int64_t *coarse_rows, double *complexity);
int grmp_amg_destroy(grmp_amg *m);

/* The linear solve of each Newton step of solve_fixpoint_full!
 * (src/solvers.jl:906-911, the cache's solve with UMFPACK) by BiCGStab,
 * right-preconditioned by one V-cycle of pc per application instead of the
 * Jacobi scaling of grmp_solve.  pc comes from grmp_amg_create on a matrix
 * with A's pattern, e.g. the first Newton matrix: its V-cycle uses the
 * values copied when it was built, so A may since have been reassembled in
 * place (the later Newton matrices of a diffusion-dominated problem).  FP64;
 * opts->initial_guess starts from x (the current iterate).  The other
 * fields and the return codes are grmp_solve's. */
int grmp_solve_bicgstab_amg(grmp_matrix *A, const grmp_vector *b, grmp_vector *x, const grmp_solve_opts *opts,
                            grmp_amg *pc, grmp_solve_info *info);

/* Anderson acceleration of the fixpoint loop: AndersonAccelerationManager
 * (src/solvers.jl:688-746) and update_anderson! (:748-808), called after
 * each linear solve of solve_fixpoint_full! (:923-929).
 * grmp_anderson_create: depth = anderson_iterations (1..10), damping =
 * anderson_damping in (0, 1], metric = the FP64-assembled norm matrix of the
 * accelerated unknowns (mass matrix for "L2", gradient stiffness for "H1";
 * NULL for "l2", the identity), ranges = nranges pairs [lo, hi) of entries
 * of the accelerated unknowns (their blocks' offsets: anderson_unknowns),
 * initial = the iterate the loop starts from (every stored iterate starts as
 * it).  grmp_anderson_update takes the fixpoint output in target and
 * overwrites its accelerated entries with the combination of stored
 * iterates; from the third call on, alpha_out (depth + 1 doubles, may be
 * NULL) gets the coefficients and *local_residual the residual of the small
 * system (the reference warns above 1e-4); before that, zeros.
 * GRMP_ERR_NONFINITE if the small system is singular or not finite.  The
 * handle keeps 2 (depth + 1) iterates (plus depth + 1 work vectors, twice with
 * a metric) of length n on the device and reads metric at every update. */
typedef struct grmp_anderson grmp_anderson;
int grmp_anderson_create(grmp_ctx *ctx, int64_t n, int depth, double damping, const grmp_matrix *metric,
                         int64_t nranges, const int64_t *ranges, const grmp_vector *initial, grmp_anderson **out);
int grmp_anderson_update(grmp_anderson *aa, grmp_vector *target, double *alpha_out, double *local_residual);
/* updates made so far minus one (AAM.cdepth of the reference: -1 after create) */
int grmp_anderson_depth(const grmp_anderson *aa);
int grmp_anderson_destroy(grmp_anderson *aa);

/* Solver plug-in: solves LinearProblem(A.entries.cscmatrix, b.entries) as the
 * linsolver of solve! (src/solvers.jl:55, 89-101, 656-660).  colptr[n+1],
 * rowval[nnz] 1-based (Int64, as SparseMatrixCSC{Float64,Int64}); fixed[nfixed]
 * 1-based penalized dofs, eliminated (x_d = b_d / A_dd).  A must be symmetric
 * positive definite on the free dofs.  precision: GRMP_PRECISION_*. */
int grmp_solve_csc(grmp_ctx *ctx, int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                   const double *b, double *x, int64_t nfixed, const int64_t *fixed, int precision, double rtol);

/* ---------------------------------------------------------------- serving */
#define GRMP_REQ_POISSON 1      /* PoissonProblem, per-cell log-normal kappa, f = 1 */
#define GRMP_REQ_ELASTICITY 2   /* LinearElasticityProblem P1{2}, Hooke scaled per cell */

typedef struct grmp_request {
    int kind;                 /* GRMP_REQ_* */
    int fe;                   /* GRMP_FE_H1P1 / GRMP_FE_H1P2 */
    int N;                    /* N x N unit square */
    uint64_t seed;            /* coefficient field seed */
    double sigma;             /* log-normal sigma */
    double E, nu;             /* elasticity */
    double rtol;
    double *x_out;            /* host buffer of ndofs doubles, filled on completion (may be NULL) */
    int precision;            /* 0 or GRMP_PRECISION_FP32_IR: FP32 sweeps with FP64 refinement and
                                 an FP64 fallback; GRMP_PRECISION_FP64: FP64 CG in the engine */
    int pad_;
} grmp_request;

typedef struct grmp_result {
    int status;
    int64_t ndofs;
    int64_t iterations;
    int outer;
    double relres;
    double t_submit, t_admit, t_done; /* seconds on the server clock */
    int fp64_fallback;        /* 1 when the FP32 sweeps stalled and FP64 CG finished the solve */
    int pad_;
} grmp_result;

typedef struct grmp_server_config {
    int mode;                 /* 1 = batched + FP32_IR (served), 2 = plain one-at-a-time FP64,
                                 3 = multigrid-preconditioned CG in FP64, mg_slots requests at once */
    int64_t memory_budget;    /* bytes of HBM for resident problems */
    int max_active;           /* problems in flight */
    int graph;                /* capture the Krylov step (mode 1) or a request's chunk of CG
                                 iterations (mode 3) as a hipGraph */
    int profile;              /* time the engine SpMV kernel with HIP events (mode 1); mode 3: the
                                 grid residual kernel (of every 4th request when graph is on) */
    int tile_rows;            /* engine rows per work tile (0 = 2048; 64 .. 8192) */
    int fp64_first_sweep;     /* > 0: a request whose first FP32 sweep takes more iterations
                                 continues in FP64 (0 = only on stalls) */
    int mg_slots;             /* mode 3: requests solved concurrently, one HIP stream each (0 = 8) */
    int mg_threads;           /* mode 3: host threads launching for those slots (0 = 4), or with
                                 mg_batch the threads that set requests up (0 = 2; -1: the batch
                                 thread sets them up itself, one HIP-calling thread) */
    int mg_block_rows;        /* mode 3: a request with at most this many rows is solved inside one
                                 workgroup (0 = 2048) */
    int mg_coarse_rows;       /* mode 3: larger requests run their multigrid levels with at most this
                                 many rows inside one workgroup (0 = 2048) */
    int mg_batch;             /* mode 3: > 0 (default 256; pass -1 for the default) advances up to this
                                 many requests in lockstep, one launch per phase over all of them, with
                                 mg_threads setup threads feeding it; 0 = one stream per request
                                 (mg_slots, mg_threads) */
    int pad_;
} grmp_server_config;

typedef struct grmp_kernel_stats {
    double spmv_seconds;      /* summed durations of the timed SpMV launches (HIP events): the engine
                                 SpMV (mode 1), the multigrid residual on levels of >= 2^21 rows (mode 3) */
    int64_t spmv_launches;
    double spmv_bytes;        /* algorithmic bytes of those launches (engine: nnz*(4+4) + 4(n+1) + 4n (x)
                                 + 4n (y) per row tile; multigrid residual: nnz*(8+4) + 4(n+1) + 8n (x)
                                 + 8n (b) + 8n (r) + n (fixed mask)) */
    int64_t steps;            /* engine steps run */
    int64_t templates;        /* problem shapes built (mesh + dofmap + pattern) */
    double template_seconds;  /* host-observed time spent building them */
} grmp_kernel_stats;

/* Request loop (continuous batching): requests are admitted by HBM footprint
 * as memory and slots free up, assembled by batched kernels and solved by a
 * block-diagonal Krylov step shared by every problem in flight.  Runs its own
 * host thread and HIP stream. */
int grmp_server_create(grmp_ctx *ctx, const grmp_server_config *cfg, grmp_server **out);
int grmp_server_submit(grmp_server *s, const grmp_request *req, int64_t *ticket);
/* GRMP_OK when done (result filled), GRMP_ERR_BUSY while pending */
int grmp_server_poll(grmp_server *s, int64_t ticket, grmp_result *res);
int grmp_server_wait(grmp_server *s, int64_t ticket, grmp_result *res);
double grmp_server_clock(const grmp_server *s);
/* engine counters since creation (profile = 1 for the SpMV timings) */
int grmp_server_stats(grmp_server *s, grmp_kernel_stats *out);
int grmp_server_destroy(grmp_server *s);
/* footprint in bytes the server would reserve for a request */
int64_t grmp_request_footprint(const grmp_request *req);
int64_t grmp_request_ndofs(const grmp_request *req);

#ifdef __cplusplus
}
#endif
#endif /* GRMP_H */
