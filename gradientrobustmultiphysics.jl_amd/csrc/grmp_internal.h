// Internal declarations of libgrmp_hip.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/grmp.h"

// ------------------------------------------------------------------ tuning
// Launch shape of the element assembly of one form kind: the kernel variant
// (1: one thread walks a row's cells; 2: the block's cell work staged in LDS
// over all threads; 3: as 2 with each thread's cell gathers batched), the
// rows one workgroup owns, and whether consecutive row
// blocks go to one XCD (1) or round-robin over the XCDs (0, the hardware's
// dispatch order).  Zero fields mean the built-in choice.  Results do not
// depend on it: every entry is summed in the same order by every variant.
struct AsmTune {
    int variant = 0;
    int rows = 0;
    int remap = -1;   // -1: built-in
};
// tuning table entry, keyed by (form kind, quadrature points, nnz/row bucket)
struct TuneEntry {
    int kind = 0, nq = 0, bucket = 0;
    AsmTune t;
};
int nnz_row_bucket(int max_row);  // 0: <= 8, 1: <= 16, 2: <= 32, 3: <= 64, 4: above

// ------------------------------------------------------------------ errors
struct grmp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::mutex mu;
    // reusable device scratch for library primitives (scan temp storage, flags)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int *dflag = nullptr;  // device error/overflow flag
    std::vector<TuneEntry> tune;  // assembly launch shapes (tuning table)
    int64_t max_nnz = 2147483647; // pattern builds above this many nonzeros fail (<= the int32 CSR limit)
};

int grmp_fail(grmp_ctx *ctx, int code, const char *fmt, ...);

#define GRMP_HIP(ctx, call)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return grmp_fail((ctx), e_ == hipErrorOutOfMemory ? GRMP_ERR_OOM : GRMP_ERR_HIP, \
                             "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                             __FILE__, __LINE__);                                        \
    } while (0)

#define GRMP_CHECK(ctx, cond, code, ...)                 \
    do {                                                 \
        if (!(cond)) return grmp_fail((ctx), (code), __VA_ARGS__); \
    } while (0)

#define GRMP_TRY(expr)            \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != GRMP_OK) return rc_; \
    } while (0)

// device allocation helpers (stream-ordered)
template <class T>
int dalloc(grmp_ctx *ctx, T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    GRMP_HIP(ctx, hipMallocAsync((void **)p, count * sizeof(T), ctx->stream));
    return GRMP_OK;
}
template <class T>
void dfree(grmp_ctx *ctx, T *&p)
{
    if (p) (void)hipFreeAsync((void *)p, ctx->stream);
    p = nullptr;
}
int ensure_scratch(grmp_ctx *ctx, size_t bytes);

// ------------------------------------------------------------------ objects
struct grmp_vector {
    grmp_ctx *ctx;
    int64_t n;
    double *d;
};

enum MeshKind { MESH_GENERIC = 0, MESH_UNITSQUARE = 1, MESH_UNITCUBE = 2 };

struct grmp_mesh {
    grmp_ctx *ctx;
    int dim;
    int64_t nnodes, ncells;
    int nloc;            // vertices per cell
    double *coords;      // SoA [dim][nnodes]
    int32_t *cells;      // SoA [nloc][ncells], 0-based
    int kind;            // MeshKind
    int N;               // structured resolution
    int8_t *fsign;       // SoA [3][ncells] CellFaceSigns (triangles; nullptr if unknown)
    // faces of uploaded triangle meshes, numbered by first appearance in cell
    // order (generated unit squares number theirs in closed form instead)
    int32_t *cellfaces;  // SoA [3][ncells], or nullptr
    uint8_t *bface;      // [nfaces] 1 for a boundary face (one cell)
    int64_t nfaces;
};

struct grmp_space {
    grmp_ctx *ctx;
    const grmp_mesh *mesh;
    int fe, ncomp;
    int64_t ndofs;
    int nloc;            // dofs per cell (all components)
    int32_t *celldofs;   // SoA [nloc][ncells], 0-based
    int32_t *inc_ptr;    // [ndofs+1] cell incidences per dof
    int32_t *inc;        // [ncells*nloc] packed cell*nloc + local, sorted per dof
    int max_inc;
    int64_t nfaces;      // P2 on unit square: number of faces
};

// element/form kinds resolved on the device
enum FormKind {
    FK_LAPLACE_P1TRI = 1,
    FK_LAPLACE_P2TRI = 2,
    FK_LAPLACE_P1TET = 3,
    FK_HOOKE_P1VTRI = 4,
    FK_MASS_P1TRI = 5,
    FK_MASS_P2TRI = 6,
    FK_LAPLACE_BR = 7,   // (nu grad u, grad v) on H1BR{2}
    FK_DIV_BRP0 = 8,     // LagrangeMultiplier(Divergence): rows BR, columns P0
    FK_DIVT_P0BR = 9,    // its transposed copy: rows P0, columns BR
    FK_NLDIFF_P1TRI = 10, // Newton matrix of alpha(u) grad u on P1 triangles
    FK_NLDIFF_P2TRI = 11, // ... on P2 triangles
};

struct DevForm {
    int kind;
    int nq;              // quadrature points
    int coef_cell;       // kappa through the action
    int rnloc, cnloc;    // local dofs of row / column space
    int words;           // slot words per incidence
    int pad_;
    double factor;
    double factor_t;     // transposed-copy factor (FK_DIVT_P0BR)
    double lam, mu;
    const double *kappa;
    const int8_t *fsign; // CellFaceSigns of the mesh (H1BR)
    const double *coords;
    const int32_t *cells;
    int64_t nnodes, ncells;
    const int32_t *rdofs, *inc_ptr, *inc;
    const int32_t *cdofs;
    int64_t ro, co;      // global offsets of row / column block
    uint32_t *pos;       // slot map (uint8 positions relative to the row start)
    // Newton forms: the iterate at this space's dofs (nullptr while the
    // pattern is built: then every local pair counts), the quadrature table
    // (x0, x1, w per point) and alpha's coefficients
    const double *uold;
    const double *qtab;
    int nqt, pad2_;
    double nlc[3];
};

struct FormRec {
    grmp_form f;
    const grmp_space *rs, *cs;
    int kind;
    int nq;
    uint32_t *pos = nullptr;
    int words = 0;
    double *qtab = nullptr;  // Newton forms: device quadrature table
    int nqt = 0;
};

struct grmp_matrix {
    grmp_ctx *ctx;
    int nblocks;
    const grmp_space *spaces[8];
    int64_t offs[9];
    int64_t n;
    std::vector<FormRec> forms;
    // pattern
    bool have_pattern = false;
    int64_t nnz = 0;
    int max_row = 0;
    int32_t *rowptr = nullptr;  // [n+1]
    int32_t *cols = nullptr;    // [nnz]
    uint8_t *diagpos = nullptr; // [n]
    uint8_t *fixed = nullptr;   // [n] mask
    std::vector<int64_t> fixed_host;
    double penalty = 1e60;
    // values
    int values = 0;
    double *val64 = nullptr;
    float *val32 = nullptr;
    double *dinv64 = nullptr;
    float *dinv32 = nullptr;
    float last_asm_ms = 0.f;
    DevForm *dforms = nullptr;      // device copy of the forms, grouped by block row
    std::vector<int> form_first;    // first form of each block row in dforms
    bool tune_set = false;          // a launch shape set on this matrix overrides the table
    AsmTune tune;
};

// ------------------------------------------------------------- internals
int build_incidence(grmp_ctx *ctx, grmp_space *s);
int exclusive_scan_i32(grmp_ctx *ctx, const int32_t *in, int32_t *out, int64_t n);  // out[n+1]
int sum_i32_as_i64(grmp_ctx *ctx, const int32_t *in, int64_t n, int64_t *total);   // host total, synchronizes
// stable radix sorts of (key, value) pairs on the low `bits` key bits, and the
// sums of runs of equal keys (host count, synchronizes); context stream
int sort_pairs_u64_f64(grmp_ctx *ctx, const uint64_t *kin, uint64_t *kout, const double *vin, double *vout,
                       int64_t n, int bits);
int sort_pairs_i32_i32(grmp_ctx *ctx, const int32_t *kin, int32_t *kout, const int32_t *vin, int32_t *vout,
                       int64_t n, int bits);
int reduce_runs_u64_f64(grmp_ctx *ctx, const uint64_t *keys, const double *vals, int64_t n, uint64_t *ukeys,
                        double *sums, int64_t *count);
// largest nonzero count of a CSR pattern with int32 row pointers and columns
constexpr int64_t GRMP_MAX_NNZ = 2147483647;
int form_resolve(grmp_matrix *A, FormRec &r);
void fill_devform(const grmp_matrix *A, const FormRec &r, DevForm &d);

// one matrix of a batched assembly (assemble_batch): a single-form matrix with
// its pattern, the cell coefficient replacing its form's (or nullptr), and
// where its FP64 values and inverse diagonals go
struct AsmBatchItem {
    const grmp_matrix *A;
    const double *kappa;
    double *val64;
    double *dinv64;
    float *dinv32;
};
// Assembles every item in one launch of a persistent kernel over a queue of
// (item, row block) tiles; values are bitwise those of per-matrix assembly.
// The staging is synchronous: the call returns after the upload.
int assemble_batch(grmp_ctx *ctx, hipStream_t st, const AsmBatchItem *items, int n);

// assembly internals shared with the server
int assemble_values(const grmp_matrix *A, const double *kappa, DevForm *dforms, DevForm *hforms, double *val64,
                    float *val32, double *dinv64, float *dinv32, hipStream_t st);
int linearform_launch(grmp_ctx *ctx, const grmp_space *s, const grmp_lform *form, double *bp, hipStream_t st);

// solver internals shared with the server
struct ProblemDesc;
int spmv_launch(grmp_ctx *ctx, const grmp_matrix *A, int values, const double *x, double *y,
                hipStream_t st);
// x = b/A_dd on fixed rows, 0 elsewhere
int init_x_launch(const grmp_matrix *A, const double *val64, const double *b, double *x, hipStream_t st);
// b[i] = 0 on fixed rows (homogeneous Dirichlet, penalty * 0)
int zero_fixed_launch(const grmp_matrix *A, double *b, hipStream_t st);
int plain_cg_solve(grmp_matrix *A, const double *b, double *x, const grmp_solve_opts &o, grmp_solve_info &info);
// vel_mg: the multigrid block of the block preconditioner (nullptr: diagonal)
int minres_solve(grmp_matrix *A, const double *bd, double *xd, const grmp_solve_opts &o, grmp_solve_info &info,
                 grmp_mg *vel_mg = nullptr, int vel_ncomp = 0);

// multigrid-preconditioned CG on unit-square hierarchies (grmp_mg.hip)
struct MgLevelIn {
    const grmp_matrix *A;    // pattern of the level (values are the job's own)
    const grmp_space *sp;
    int N, fe, ncomp;
};
struct MgStatus {            // host copy of the job's scalars
    double relres;
    int it, done, status, pad;
};
struct MgJob;
int64_t mg_job_bytes(const MgLevelIn *in, int nlev);
// block_rows: a request with at most this many rows runs whole inside one
// workgroup; coarse_rows: otherwise its levels with at most this many rows do;
// graph: capture the grid-path chunk of iterations once and replay it
int mg_job_create(grmp_ctx *ctx, hipStream_t st, const MgLevelIn *in, int nlev, bool profile, int block_rows,
                  int coarse_rows, bool graph, MgJob **out);
double *mg_job_kappa0(MgJob *j);     // level-0 cell coefficient, filled by the caller
double *mg_job_rhs(MgJob *j);        // level-0 right-hand side, filled by the caller
const double *mg_job_solution(const MgJob *j);
int mg_job_setup(MgJob *j, double rtol, int maxit);
// mg_job_setup for n jobs on one stream with their level assembly in one launch
int mg_jobs_setup_batched(MgJob *const *jobs, int n, hipStream_t st, double *const rtol, int maxit);
// one V-cycle of a created hierarchy, z = V v on the hierarchy's stream (the
// velocity block of the Stokes preconditioner); level-0 rows of the hierarchy
int mg_precond_apply(grmp_mg *m, const double *v, double *z, hipStream_t st);
// one V-cycle of an algebraic multigrid hierarchy as a preconditioner
int amg_precond_apply(grmp_amg *m, const double *v, double *z, hipStream_t st);
int64_t amg_level0_rows(const grmp_amg *m);
grmp_ctx *amg_context(const grmp_amg *m);
// two components at once: a second copy of the hierarchy on its own stream
int mg_pair_prepare(grmp_mg *m);
int mg_precond_apply_pair(grmp_mg *m, const double *v0, double *z0, const double *v1, double *z1, hipStream_t st);
int64_t mg_level0_rows(const grmp_mg *m);
int mg_job_iterate(MgJob *j, int iters);
int mg_job_verify(MgJob *j);         // after done == 2: check the true residual
bool mg_job_block(const MgJob *j);   // whole iterations in one workgroup (checks in place)
void mg_job_set_stream(MgJob *j, hipStream_t st);
int mg_job_status_async(MgJob *j, MgStatus *host);
bool mg_job_status_ready(MgJob *j);
int mg_job_wait(MgJob *j);
void mg_job_collect(MgJob *j);       // after a status event: add the finished profiled launches
void mg_job_profile(MgJob *j, double *seconds, double *bytes, int64_t *launches);
void mg_job_destroy(MgJob *j);
// lockstep batches of jobs set up on one stream (grmp_mg.hip)
struct MgBatch;
int mg_batch_create(grmp_ctx *ctx, hipStream_t st, bool profile, MgBatch **out);
int mg_batch_run(MgBatch *b, MgJob *const *jobs, int n, int iters);  // + status copy and event
int mg_batch_wait(MgBatch *b);
const MgStatus *mg_batch_status(const MgBatch *b);                  // per job, after the wait
void mg_batch_profile(const MgBatch *b, double *seconds, double *bytes, int64_t *launches);
void mg_batch_destroy(MgBatch *b);
// timing
double wall_seconds();
