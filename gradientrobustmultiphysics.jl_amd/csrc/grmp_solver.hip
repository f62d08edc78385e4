// Sparse solves of libgrmp_hip.so: CSR SpMV, the plain one-problem FP64
// Jacobi-PCG and the batched FP32-inner / FP64-refinement Krylov engine.
//
// These replace the direct factorization behind solve_direct!
// (src/solvers.jl:622-686: LinearSolve + UMFPACK at :55,656-660).  Fixed
// (penalized) dofs are eliminated instead of carried with the 1e60 penalty:
// x_fixed = b_fixed / A_dd, their rows are dropped from the Krylov space and
// the residual is measured on the free rows only, the convention of the
// reference's residual check (src/solvers.jl:663-668).
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>

#include "grmp_engine.h"

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#include "grmp_sparse.h"

// --------------------------------------------------------------- SpMV API
__global__ __launch_bounds__(KB) void k_spmv(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                            const double *__restrict__ v, const double *__restrict__ x,
                                            double *__restrict__ y, int64_t n)
{
    __shared__ double lds[SPMV_CAP64];
    for (int64_t r0 = (int64_t)blockIdx.x * KB; r0 < n; r0 += (int64_t)gridDim.x * KB) {
        const int64_t r1 = r0 + KB < n ? r0 + KB : n;
        const double s = csr_rows_lds64(rp, ci, v, x, r0, r1, lds);
        if (r0 + threadIdx.x < r1) y[r0 + threadIdx.x] = s;
    }
}

// FP32 values, FP64 input, sums and output
__global__ __launch_bounds__(KB) void k_spmv32(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                              const float *__restrict__ v, const double *__restrict__ x,
                                              double *__restrict__ y, int64_t n)
{
    __shared__ double lds[SPMV_CAP64];
    for (int64_t r0 = (int64_t)blockIdx.x * KB; r0 < n; r0 += (int64_t)gridDim.x * KB) {
        const int64_t r1 = r0 + KB < n ? r0 + KB : n;
        const double s = csr_rows_lds64(rp, ci, v, x, r0, r1, lds);
        if (r0 + threadIdx.x < r1) y[r0 + threadIdx.x] = s;
    }
}

int spmv_launch(grmp_ctx *ctx, const grmp_matrix *A, int values, const double *x, double *y, hipStream_t st)
{
    const int64_t n = A->n;
    dim3 grid((unsigned)std::min<int64_t>(nblk(n, KB), 1 << 20)), blk(KB);
    if (values == GRMP_VALUES_FP64)
        hipLaunchKernelGGL(k_spmv, grid, blk, 0, st, A->rowptr, A->cols, A->val64, x, y, n);
    else
        hipLaunchKernelGGL(k_spmv32, grid, blk, 0, st, A->rowptr, A->cols, A->val32, x, y, n);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

extern "C" int grmp_matrix_spmv(grmp_matrix *A, int values, const grmp_vector *x, grmp_vector *y)
{
    if (!A || !x || !y) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, x->n == A->n && y->n == A->n, GRMP_ERR_ARG, "grmp_matrix_spmv: vector length != %lld", (long long)A->n);
    GRMP_CHECK(ctx, (values == GRMP_VALUES_FP64 && A->val64) || (values == GRMP_VALUES_FP32 && A->val32), GRMP_ERR_ARG,
               "grmp_matrix_spmv: requested values not assembled");
    std::lock_guard<std::mutex> g(ctx->mu);
    return spmv_launch(ctx, A, values, x->d, y->d, ctx->stream);
}

// ------------------------------------------- backward Euler right-hand side
// b_i = penalty * u_i on the fixed rows of the system matrix, else
// f_i + c (M u)_i: one pass over M's rows with u gathered through LDS.
// b may be f (row i reads f_i before its thread writes b_i); u is gathered
// by other rows and must be a different buffer.
__global__ __launch_bounds__(KB) void k_timestep_rhs(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ v, const uint8_t *__restrict__ fixed,
                                                    double penalty, double c, const double *f,
                                                    const double *__restrict__ u, double *b, int64_t n)
{
    __shared__ double lds[SPMV_CAP64];
    for (int64_t r0 = (int64_t)blockIdx.x * KB; r0 < n; r0 += (int64_t)gridDim.x * KB) {
        const int64_t r1 = r0 + KB < n ? r0 + KB : n;
        const double s = csr_rows_lds64(rp, ci, v, u, r0, r1, lds);
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) b[i] = fixed[i] ? penalty * u[i] : f[i] + c * s;
    }
}

extern "C" int grmp_timestep_rhs(const grmp_matrix *S, const grmp_matrix *M, const grmp_vector *f,
                                 const grmp_vector *u, double inv_dt, grmp_vector *b)
{
    if (!S || !M || !f || !u || !b) return GRMP_ERR_ARG;
    grmp_ctx *ctx = S->ctx;
    GRMP_CHECK(ctx, M->ctx == ctx && f->ctx == ctx && u->ctx == ctx && b->ctx == ctx, GRMP_ERR_ARG,
               "grmp_timestep_rhs: objects of different contexts");
    GRMP_CHECK(ctx, S->have_pattern && M->have_pattern && M->val64, GRMP_ERR_ARG,
               "grmp_timestep_rhs: system pattern and FP64 mass matrix values needed");
    GRMP_CHECK(ctx, M->n == S->n && f->n == S->n && u->n == S->n && b->n == S->n, GRMP_ERR_ARG,
               "grmp_timestep_rhs: lengths differ from the system's %lld", (long long)S->n);
    GRMP_CHECK(ctx, b->d != u->d, GRMP_ERR_ARG, "grmp_timestep_rhs: b must not alias u");
    GRMP_CHECK(ctx, std::isfinite(inv_dt), GRMP_ERR_ARG, "grmp_timestep_rhs: 1/dt is not finite");
    std::lock_guard<std::mutex> g(ctx->mu);
    const int64_t n = S->n;
    dim3 grid((unsigned)std::min<int64_t>(nblk(n, KB), 1 << 20)), blk(KB);
    hipLaunchKernelGGL(k_timestep_rhs, grid, blk, 0, ctx->stream, M->rowptr, M->cols, M->val64, S->fixed, S->penalty,
                       inv_dt, f->d, u->d, b->d, n);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

// ------------------------------------------------------- plain FP64 PCG
// One problem, host-driven: every iteration launches its kernels and reads
// the dot products back to the host, which computes alpha/beta and checks
// convergence.  This is the conventional GPU CG and the baseline the
// batched engine is measured against.
#define PB 256
#define PGRID 1024        // grid of the streaming vector kernels
#define PGRID_ROWS 16384  // grid cap of the row kernels: enough workgroups to fill every CU slot

__global__ void k_plain_init(const int32_t *rp, const uint8_t *dpos, const double *val, const uint8_t *fixed,
                             const double *b, double *x, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fixed[i]) {
        double d = dpos[i] != 255 ? val[rp[i] + dpos[i]] : 1.0;
        x[i] = b[i] / d;
    } else {
        x[i] = 0.0;
    }
}

// r = b - A x on free rows (0 on fixed rows); z = dinv r; p = z;
// partials of (r.z, r.r)
__global__ __launch_bounds__(PB) void k_plain_resid(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                   const double *__restrict__ v, const uint8_t *__restrict__ fixed,
                                                   const double *__restrict__ dinv, const double *__restrict__ b,
                                                   const double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, double *__restrict__ p, int64_t n,
                                                   double *__restrict__ part)
{
    __shared__ double lds[SPMV_CAP64];
    double a0 = 0.0, a1 = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * PB; r0 < n; r0 += (int64_t)gridDim.x * PB) {
        const int64_t r1 = r0 + PB < n ? r0 + PB : n;
        const double s = csr_rows_lds64(rp, ci, v, x, r0, r1, lds);
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) {
            double ri = fixed[i] ? 0.0 : b[i] - s;
            double zi = dinv[i] * ri;
            r[i] = ri;
            z[i] = zi;
            p[i] = zi;
            a0 += ri * zi;
            a1 += ri * ri;
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

// Scalars of the plain CG on the device: alpha and beta are formed there
// and the convergence test runs there, so the host reads the state back
// once per chunk of iterations instead of twice per iteration.  Kernels of a
// chunk launched after convergence return at once (done is uniform).
struct PlainScal {
    double rz, rr, bb, rtol2, alpha, beta;
    long long it, maxit;
    int done, status;
};

// q = A p on free rows; partial p.q
__global__ __launch_bounds__(PB) void k_plain_spmv(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const double *__restrict__ v, const uint8_t *__restrict__ fixed,
                                                  const double *__restrict__ p, double *__restrict__ q, int64_t n,
                                                  double *__restrict__ part, const PlainScal *__restrict__ S)
{
    if (S->done) return;
    __shared__ double lds[SPMV_CAP64];
    double a0 = 0.0, a1 = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * PB; r0 < n; r0 += (int64_t)gridDim.x * PB) {
        const int64_t r1 = r0 + PB < n ? r0 + PB : n;
        const double s = csr_rows_lds64(rp, ci, v, p, r0, r1, lds);
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) {
            double qi = fixed[i] ? 0.0 : s;
            q[i] = qi;
            a0 += p[i] * qi;
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

// x += alpha p; r -= alpha q; z = dinv r; partials (r.z, r.r)
__global__ void k_plain_axpy(const PlainScal *__restrict__ S, const double *__restrict__ p,
                             const double *__restrict__ q, const double *__restrict__ dinv, double *__restrict__ x,
                             double *__restrict__ r, double *__restrict__ z, int64_t n, double *__restrict__ part)
{
    if (S->done) return;
    const double alpha = S->alpha;
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * q[i];
        double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        a0 += ri * zi;
        a1 += ri * ri;
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

__global__ void k_plain_pupd(const PlainScal *__restrict__ S, const double *__restrict__ z, double *__restrict__ p,
                             int64_t n)
{
    if (S->done) return;
    const double beta = S->beta;
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) p[i] = z[i] + beta * p[i];
}

// alpha = rz / (p.q) from the product's partials
__global__ __launch_bounds__(PB) void k_plain_alpha(const double *__restrict__ part, int grid, PlainScal *S)
{
    if (S->done) return;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < grid; k += PB) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) S->alpha = S->rz / a;
}

// beta = rz' / rz, the iteration count and the tests the host loop made
// before each iteration: non-finite, converged, iteration cap
__global__ __launch_bounds__(PB) void k_plain_beta(const double *__restrict__ part, int grid, PlainScal *S)
{
    if (S->done) return;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < grid; k += PB) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        S->beta = a / S->rz;
        S->rz = a;
        S->rr = b;
        S->it += 1;
        if (!(b == b) || isinf(b)) {
            S->done = 1;
            S->status = GRMP_ERR_NONFINITE;
        } else if (b <= S->rtol2 * S->bb || b == 0.0) {
            S->done = 1;
        } else if (S->it >= S->maxit) {
            S->done = 1;
            S->status = GRMP_ERR_NOCONV;
        }
    }
}

// Sums a pass's per-workgroup partial pairs in a fixed order (thread t
// takes pairs t, t + PB, ..., then the block tree) into out[0..1].
__global__ __launch_bounds__(PB) void k_plain_fold(const double *__restrict__ part, int grid, double *__restrict__ out)
{
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < grid; k += PB) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// folds the partials on the device and reads the pair back
static int plain_sum(grmp_ctx *ctx, const double *dpart, double *dout, double *hpart, int grid, double &a, double &b)
{
    hipLaunchKernelGGL(k_plain_fold, dim3(1), dim3(PB), 0, ctx->stream, dpart, grid, dout);
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_HIP(ctx, hipMemcpyAsync(hpart, dout, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    a = hpart[0];
    b = hpart[1];
    return GRMP_OK;
}

int plain_cg_solve(grmp_matrix *A, const double *bd, double *xd, const grmp_solve_opts &o, grmp_solve_info &info)
{
    grmp_ctx *ctx = A->ctx;
    const int64_t n = A->n;
    GRMP_CHECK(ctx, A->rowptr && A->cols && A->val64 && A->dinv64 && A->fixed && A->diagpos && bd && xd, GRMP_ERR_ARG,
               "plain CG: matrix not assembled in FP64");
    // every exit path releases the work vectors and the scalar records
    struct Bufs {
        grmp_ctx *ctx;
        double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *part = nullptr;
        PlainScal *dsc = nullptr, *hsc = nullptr;
        ~Bufs()
        {
            if (hsc) (void)hipHostFree(hsc);
            dfree(ctx, r);
            dfree(ctx, z);
            dfree(ctx, p);
            dfree(ctx, q);
            dfree(ctx, part);
            dfree(ctx, dsc);
        }
    } B{ctx};
    GRMP_TRY(dalloc(ctx, &B.r, n));
    GRMP_TRY(dalloc(ctx, &B.z, n));
    GRMP_TRY(dalloc(ctx, &B.p, n));
    GRMP_TRY(dalloc(ctx, &B.q, n));
    GRMP_TRY(dalloc(ctx, &B.part, 2 * PGRID_ROWS + 2));
    GRMP_TRY(dalloc(ctx, &B.dsc, 1));
    GRMP_HIP(ctx, hipHostMalloc((void **)&B.hsc, sizeof(PlainScal), hipHostMallocDefault));
    double *r = B.r, *z = B.z, *p = B.p, *q = B.q, *part = B.part;
    PlainScal *dsc = B.dsc, *hsc = B.hsc;
    double *fold = part + 2 * PGRID_ROWS;
    double *hpart = &hsc->alpha;  // two doubles for the first fold
    const int grid = (int)std::min<int64_t>(PGRID_ROWS, std::max<int64_t>(1, (n + PB - 1) / PB));
    const int gride = (int)std::min<int64_t>(PGRID, grid);
    hipLaunchKernelGGL(k_plain_init, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, A->rowptr, A->diagpos, A->val64, A->fixed,
                       bd, xd, n);
    hipLaunchKernelGGL(k_plain_resid, dim3(grid), dim3(PB), 0, ctx->stream, A->rowptr, A->cols, A->val64, A->fixed,
                       A->dinv64, bd, xd, r, z, p, n, part);
    double rz, rr;
    int rc = plain_sum(ctx, part, fold, hpart, grid, rz, rr);
    const double bb = rr;
    int64_t it = 0;
    int status = GRMP_OK;
    // the tests of the first iteration on the host, the rest on the device
    if (rc == GRMP_OK) {
        if (!(rr == rr) || std::isinf(rr)) status = GRMP_ERR_NONFINITE;
        else if (rr <= o.rtol * o.rtol * bb || rr == 0.0) status = GRMP_OK;
        else if (it >= o.maxit) status = GRMP_ERR_NOCONV;
        else {
            PlainScal s0;
            memset(&s0, 0, sizeof(s0));
            s0.rz = rz;
            s0.rr = rr;
            s0.bb = bb;
            s0.rtol2 = o.rtol * o.rtol;
            s0.maxit = o.maxit;
            *hsc = s0;
            if (hipMemcpyAsync(dsc, hsc, sizeof(PlainScal), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
                rc = grmp_fail(ctx, GRMP_ERR_HIP, "plain CG: scalars");
            const int chunk = 8;
            while (rc == GRMP_OK) {
                for (int k = 0; k < chunk; ++k) {
                    hipLaunchKernelGGL(k_plain_spmv, dim3(grid), dim3(PB), 0, ctx->stream, A->rowptr, A->cols,
                                       A->val64, A->fixed, p, q, n, part, dsc);
                    hipLaunchKernelGGL(k_plain_alpha, dim3(1), dim3(PB), 0, ctx->stream, part, grid, dsc);
                    hipLaunchKernelGGL(k_plain_axpy, dim3(gride), dim3(PB), 0, ctx->stream, dsc, p, q, A->dinv64, xd,
                                       r, z, n, part);
                    hipLaunchKernelGGL(k_plain_beta, dim3(1), dim3(PB), 0, ctx->stream, part, gride, dsc);
                    hipLaunchKernelGGL(k_plain_pupd, dim3(gride), dim3(PB), 0, ctx->stream, dsc, z, p, n);
                }
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(hsc, dsc, sizeof(PlainScal), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                    hipStreamSynchronize(ctx->stream) != hipSuccess) {
                    rc = grmp_fail(ctx, GRMP_ERR_HIP, "plain CG: iteration chunk");
                    break;
                }
                if (hsc->done) break;
            }
            if (rc == GRMP_OK) {
                it = hsc->it;
                rr = hsc->rr;
                status = hsc->status;
            }
        }
    }
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    info.iterations = it;
    info.outer = 1;
    info.relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    info.status = rc ? rc : status;
    return rc;
}

// ------------------------------------------------------ BiCGStab (FP64)
// For the nonsymmetric Newton matrices of NonlinearForms: BiCGStab (van der
// Vorst 1992) with right Jacobi preconditioning, fixed dofs eliminated as in
// the plain CG (x_d = b_d / A_dd, rows masked).  Host-driven: the dot
// products are folded on the device in a fixed order and read back.
// Near-breakdowns (the shadow residual nearly orthogonal to r or to A p, or
// omega = 0) restart the recurrence from the true residual with a new
// shadow residual; on Newton matrices of P2 forms at 10^5 dofs and more
// they happen every few hundred iterations.

// y = A x on free rows (0 on fixed rows); partials (w.y, y.y)
__global__ __launch_bounds__(PB) void k_bic_spmv(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                const double *__restrict__ v, const uint8_t *__restrict__ fixed,
                                                const double *__restrict__ x, const double *__restrict__ w,
                                                double *__restrict__ y, int64_t n, double *__restrict__ part)
{
    __shared__ double lds[SPMV_CAP64];
    double a0 = 0.0, a1 = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * PB; r0 < n; r0 += (int64_t)gridDim.x * PB) {
        const int64_t r1 = r0 + PB < n ? r0 + PB : n;
        const double s = csr_rows_lds64(rp, ci, v, x, r0, r1, lds);
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) {
            const double yi = fixed[i] ? 0.0 : s;
            y[i] = yi;
            a0 += w[i] * yi;
            a1 += yi * yi;
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

// r = b - A x on free rows, rh = p = r; partials (r.r, r.r)
__global__ __launch_bounds__(PB) void k_bic_resid(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                 const double *__restrict__ v, const uint8_t *__restrict__ fixed,
                                                 const double *__restrict__ b, const double *__restrict__ x,
                                                 double *__restrict__ r, double *__restrict__ rh,
                                                 double *__restrict__ p, int64_t n, double *__restrict__ part)
{
    __shared__ double lds[SPMV_CAP64];
    double a0 = 0.0, a1 = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * PB; r0 < n; r0 += (int64_t)gridDim.x * PB) {
        const int64_t r1 = r0 + PB < n ? r0 + PB : n;
        const double s = csr_rows_lds64(rp, ci, v, x, r0, r1, lds);
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) {
            const double ri = fixed[i] ? 0.0 : b[i] - s;
            r[i] = ri;
            rh[i] = ri;
            p[i] = ri;
            a0 += ri * ri;
            a1 += ri * ri;
        }
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

// x = g on free rows (the initial guess), fixed rows keep x_d = b_d / A_dd
__global__ void k_bic_guess(const uint8_t *__restrict__ fixed, const double *__restrict__ g, double *__restrict__ x,
                            int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB)
        if (!fixed[i]) x[i] = g[i];
}

__global__ void k_bic_scale(const double *__restrict__ d, const double *__restrict__ in, double *__restrict__ out,
                            int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) out[i] = d[i] * in[i];
}

// s = r - alpha v; partials (s.s, 0)
__global__ __launch_bounds__(PB) void k_bic_s(const double *__restrict__ r, const double *__restrict__ v, double alpha,
                                             double *__restrict__ s, int64_t n, double *__restrict__ part)
{
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        a0 += si * si;
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

// x += alpha ph + omega sh; r = s - omega t; partials (rh.r, r.r)
__global__ __launch_bounds__(PB) void k_bic_update(double *__restrict__ x, const double *__restrict__ ph,
                                                  const double *__restrict__ sh, double alpha, double omega,
                                                  const double *__restrict__ s, const double *__restrict__ t,
                                                  double *__restrict__ r, const double *__restrict__ rh, int64_t n,
                                                  double *__restrict__ part)
{
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) {
        x[i] += alpha * ph[i] + omega * sh[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        a0 += rh[i] * ri;
        a1 += ri * ri;
    }
    block_sum2(a0, a1);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a0;
        part[2 * blockIdx.x + 1] = a1;
    }
}

__global__ void k_bic_p(const double *__restrict__ r, const double *__restrict__ v, double beta, double omega,
                        double *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB)
        p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// pc: the right preconditioner (one algebraic multigrid V-cycle) in place of
// the Jacobi scaling, or nullptr
int bicgstab_solve(grmp_matrix *A, const double *bd, double *xd, const grmp_solve_opts &o, grmp_solve_info &info,
                   grmp_amg *pc = nullptr)
{
    grmp_ctx *ctx = A->ctx;
    const int64_t n = A->n;
    GRMP_CHECK(ctx, A->rowptr && A->cols && A->val64 && A->dinv64 && A->fixed && A->diagpos && bd && xd, GRMP_ERR_ARG,
               "BiCGStab: matrix not assembled in FP64");
    struct Bufs {
        grmp_ctx *ctx;
        double *w = nullptr;
        double *hp = nullptr;
        ~Bufs()
        {
            if (hp) (void)hipHostFree(hp);
            dfree(ctx, w);
        }
    } B{ctx};
    GRMP_TRY(dalloc(ctx, &B.w, 8 * (size_t)n + 2 * PGRID_ROWS + 2));
    GRMP_HIP(ctx, hipHostMalloc((void **)&B.hp, 2 * sizeof(double), hipHostMallocDefault));
    double *r = B.w, *rh = r + n, *p = rh + n, *v = p + n, *s = v + n, *t = s + n, *ph = t + n, *sh = ph + n;
    double *part = sh + n, *fold = part + 2 * PGRID_ROWS;
    const int grid = (int)std::min<int64_t>(PGRID_ROWS, std::max<int64_t>(1, (n + PB - 1) / PB));
    const int gride = (int)std::min<int64_t>(PGRID, grid);
    if (o.initial_guess) GRMP_HIP(ctx, hipMemcpyAsync(sh, xd, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_plain_init, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, A->rowptr, A->diagpos, A->val64,
                       A->fixed, bd, xd, n);
    hipLaunchKernelGGL(k_bic_resid, dim3(grid), dim3(PB), 0, ctx->stream, A->rowptr, A->cols, A->val64, A->fixed, bd, xd,
                       r, rh, p, n, part);
    double rho, rr;
    GRMP_TRY(plain_sum(ctx, part, fold, B.hp, grid, rho, rr));
    const double bb = rr, tol2 = o.rtol * o.rtol * bb;
    // r = rh = p = b - A x at the current x
    auto restart = [&](double &rho_, double &rr_) -> int {
        hipLaunchKernelGGL(k_bic_resid, dim3(grid), dim3(PB), 0, ctx->stream, A->rowptr, A->cols, A->val64, A->fixed, bd,
                           xd, r, rh, p, n, part);
        return plain_sum(ctx, part, fold, B.hp, grid, rho_, rr_);
    };
    if (o.initial_guess) {
        hipLaunchKernelGGL(k_bic_guess, dim3(gride), dim3(PB), 0, ctx->stream, A->fixed, sh, xd, n);
        GRMP_TRY(restart(rho, rr));
    }
    double rhh = rr;              // rh.rh
    const double brk = 1e-15;     // near-breakdown: |cosine| below this
    const int max_restarts = 200;
    int restarts = 0;
    int64_t it = 0;
    int status = GRMP_OK;
    while (true) {
        if (!(rr == rr) || std::isinf(rr)) {
            status = GRMP_ERR_NONFINITE;
            break;
        }
        if (rr <= tol2 || rr == 0.0) break;
        if (it >= o.maxit) {
            status = GRMP_ERR_NOCONV;
            break;
        }
        if (std::fabs(rho) <= brk * std::sqrt(rhh * rr)) {
            if (++restarts > max_restarts) {
                status = GRMP_ERR_NOCONV;
                break;
            }
            GRMP_TRY(restart(rho, rr));
            rhh = rr;
            continue;
        }
        double rv, vv, ss, dummy, ts, tt;
        if (pc) GRMP_TRY(amg_precond_apply(pc, p, ph, ctx->stream));
        else hipLaunchKernelGGL(k_bic_scale, dim3(gride), dim3(PB), 0, ctx->stream, A->dinv64, p, ph, n);
        hipLaunchKernelGGL(k_bic_spmv, dim3(grid), dim3(PB), 0, ctx->stream, A->rowptr, A->cols, A->val64, A->fixed, ph,
                           rh, v, n, part);
        GRMP_TRY(plain_sum(ctx, part, fold, B.hp, grid, rv, vv));
        if (std::fabs(rv) <= brk * std::sqrt(rhh * vv)) {
            if (++restarts > max_restarts) {
                status = GRMP_ERR_NOCONV;
                break;
            }
            GRMP_TRY(restart(rho, rr));
            rhh = rr;
            continue;
        }
        const double alpha = rho / rv;
        hipLaunchKernelGGL(k_bic_s, dim3(gride), dim3(PB), 0, ctx->stream, r, v, alpha, s, n, part);
        GRMP_TRY(plain_sum(ctx, part, fold, B.hp, gride, ss, dummy));
        ++it;
        if (ss <= tol2) {  // x += alpha ph; the residual is s
            hipLaunchKernelGGL(k_bic_update, dim3(gride), dim3(PB), 0, ctx->stream, xd, ph, ph, alpha, 0.0, s, s, r, rh,
                               n, part);
            GRMP_TRY(plain_sum(ctx, part, fold, B.hp, gride, dummy, rr));
            break;
        }
        if (pc) GRMP_TRY(amg_precond_apply(pc, s, sh, ctx->stream));
        else hipLaunchKernelGGL(k_bic_scale, dim3(gride), dim3(PB), 0, ctx->stream, A->dinv64, s, sh, n);
        hipLaunchKernelGGL(k_bic_spmv, dim3(grid), dim3(PB), 0, ctx->stream, A->rowptr, A->cols, A->val64, A->fixed, sh,
                           s, t, n, part);
        GRMP_TRY(plain_sum(ctx, part, fold, B.hp, grid, ts, tt));
        const double omega = tt > 0.0 ? ts / tt : 0.0;
        hipLaunchKernelGGL(k_bic_update, dim3(gride), dim3(PB), 0, ctx->stream, xd, ph, sh, alpha, omega, s, t, r, rh,
                           n, part);
        double rho2;
        GRMP_TRY(plain_sum(ctx, part, fold, B.hp, gride, rho2, rr));
        if (omega == 0.0) {   // x and r advanced by alpha only: restart from there
            if (rr <= tol2) break;
            if (++restarts > max_restarts) {
                status = GRMP_ERR_NOCONV;
                break;
            }
            GRMP_TRY(restart(rho, rr));
            rhh = rr;
            continue;
        }
        const double beta = (rho2 / rho) * (alpha / omega);
        rho = rho2;
        hipLaunchKernelGGL(k_bic_p, dim3(gride), dim3(PB), 0, ctx->stream, r, v, beta, omega, p, n);
    }
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    info.iterations = it;
    info.outer = 1 + restarts;
    info.relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    info.status = status;
    return GRMP_OK;
}

__global__ void k_zero_fixed(const uint8_t *fixed, double *b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && fixed[i]) b[i] = 0.0;
}

int init_x_launch(const grmp_matrix *A, const double *val64, const double *b, double *x, hipStream_t st)
{
    GRMP_CHECK(A->ctx, A->rowptr && A->diagpos && A->fixed && val64 && b && x, GRMP_ERR_ARG,
               "init_x_launch: pattern or FP64 values missing");
    hipLaunchKernelGGL(k_plain_init, dim3(nblk(A->n, 256)), dim3(256), 0, st, A->rowptr, A->diagpos, val64, A->fixed, b,
                       x, A->n);
    GRMP_HIP(A->ctx, hipGetLastError());
    return GRMP_OK;
}

int zero_fixed_launch(const grmp_matrix *A, double *b, hipStream_t st)
{
    GRMP_CHECK(A->ctx, A->fixed && b, GRMP_ERR_ARG, "zero_fixed_launch: fixed mask or vector missing");
    hipLaunchKernelGGL(k_zero_fixed, dim3(nblk(A->n, 256)), dim3(256), 0, st, A->fixed, b, A->n);
    GRMP_HIP(A->ctx, hipGetLastError());
    return GRMP_OK;
}

// ============================================================ the engine
//
// Step = 7 launches over the tile table:
//   spmv  (ITER)            q = A32 p, tile partials of p.q
//   red A                   alpha = rz / p.q
//   axpy  (ITER)            dx += alpha p, r -= alpha q, tile partials (r.z, r.r)
//   red B                   convergence estimate, reliable-update trigger, beta
//   elem  (ITER)            p = D r + beta p
//         (UPDATE)          x += dx / s (FP64), and p = D r + beta p
//         (INIT/REINIT)     r = s r64, dx = 0, p = D r | p *= s/s_old
//   resid (RESID)           r64 = b - A64 x, tile partials r64.r64   (FP64)
//   red C                   phase changes of elem / resid
// Tile kernels write one pair of partials per tile; the reduce kernels sum a
// problem's partials in a fixed order (one block per problem).  Kernel
// boundaries order the memory, so no fence or atomic is needed on the data
// path, and a problem's arithmetic does not depend on its neighbours.
//
// Mixed precision: CG with reliable updates (FP32 recurrences; the FP64
// residual b - A x replaces the FP32 one whenever the latter has dropped by
// delta since the last refresh, keeping the search direction), so the
// Krylov space is not discarded as in restarted refinement.

#define ENGINE_ARGS KProb *__restrict__ P, const KTile *__restrict__ T, const KRange *__restrict__ R, \
                    double *__restrict__ part, const int *__restrict__ ntiles, unsigned long long *__restrict__ nbytes

// q = A32 p over rows row0 .. row1 of one problem (FP32 products and sums),
// staged through LDS one KB-row block at a time; returns the partial p.q
__device__ __forceinline__ double spmv_tile(const KProb &pr, int row0, int row1, float *lds)
{
    double acc = 0.0;
    for (int r0 = row0; r0 < row1; r0 += KB) {
        const int r1 = r0 + KB < row1 ? r0 + KB : row1;
        const float s = csr_rows_lds<float>(pr.rowptr, pr.cols, pr.val32, pr.p, r0, r1, lds);
        const int i = r0 + (int)threadIdx.x;
        if (i < r1) {
            pr.q[i] = s;
            acc += (double)pr.p[i] * (double)s;
        }
    }
    return acc;
}

// FP64 fallback: q64 = A64 p64 on free rows; returns the partial p.q.  lds
// holds SPMV_CAP / 2 doubles (the FP32 tiles' buffer).
__device__ __forceinline__ double spmv_tile64(const KProb &pr, int row0, int row1, double *lds)
{
    double acc = 0.0;
    for (int r0 = row0; r0 < row1; r0 += KB) {
        const int r1 = r0 + KB < row1 ? r0 + KB : row1;
        const double s = csr_rows_lds<double, double, double, SPMV_CAP / 2>(pr.rowptr, pr.cols, pr.val64, pr.p64, r0,
                                                                             r1, lds);
        const int i = r0 + (int)threadIdx.x;
        if (i < r1) {
            const double qi = pr.fixed[i] ? 0.0 : s;
            pr.q64[i] = qi;
            acc += pr.p64[i] * qi;
        }
    }
    return acc;
}

__global__ __launch_bounds__(KB) void k_eng_spmv(ENGINE_ARGS)
{
    // SPMV_CAP floats; FP64 fallback tiles stage SPMV_CAP / 2 doubles in it
    __shared__ __align__(16) float lds32[SPMV_CAP];
    unsigned long long myb = 0;
    const int nt = *ntiles;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const KTile tl = T[t];
        const int pi = tl.prob;
        if (P[pi].phase != PH_ITER) continue;
        const KProb pr = P[pi];
        double acc = pr.fp64 ? spmv_tile64(pr, tl.row0, tl.row1, (double *)lds32) : spmv_tile(pr, tl.row0, tl.row1, lds32);
        double dummy = 0.0;
        block_sum2(acc, dummy);
        if (threadIdx.x == 0) {
            part[2 * t] = acc;
            // algorithmic bytes of this tile: values + columns, row pointers, x gather, y write
            const unsigned long long rows = (unsigned long long)(tl.row1 - tl.row0);
            const unsigned long long nz = (unsigned long long)(pr.rowptr[tl.row1] - pr.rowptr[tl.row0]);
            const unsigned long long sv = pr.fp64 ? 8ull : 4ull;
            myb += nz * (sv + 4ull) + 4ull * (rows + 1) + 2ull * sv * rows;
        }
    }
    // per-block byte count, summed by k_eng_bytes (no atomics in the hot loop)
    if (threadIdx.x == 0) nbytes[blockIdx.x] = myb;
}

// adds the per-block SpMV byte counts of one step to the running total
__global__ __launch_bounds__(KB) void k_eng_bytes(const unsigned long long *__restrict__ blk, int nblk,
                                                  unsigned long long *__restrict__ total)
{
    __shared__ unsigned long long sh[KB];
    unsigned long long a = 0;
    for (int k = threadIdx.x; k < nblk; k += KB) a += blk[k];
    sh[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < KB; ++k) t += sh[k];
        *total += t;
    }
}

__global__ __launch_bounds__(KB) void k_eng_axpy(ENGINE_ARGS)
{
    const int nt = *ntiles;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const KTile tl = T[t];
        const int pi = tl.prob;
        if (P[pi].phase != PH_ITER) continue;
        const KProb pr = P[pi];
        double rz = 0.0, rr = 0.0;
        if (pr.fp64) {
            // FP64 fallback: x += alpha p; r -= alpha q; z = D r
            const double a = pr.alpha;
            for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) {
                pr.x[i] += a * pr.p64[i];
                const double ri = pr.r64[i] - a * pr.q64[i];
                pr.r64[i] = ri;
                rz += ri * (pr.dinv64[i] * ri);
                rr += ri * ri;
            }
            block_sum2(rz, rr);
            if (threadIdx.x == 0) {
                part[2 * t] = rz;
                part[2 * t + 1] = rr;
            }
            continue;
        }
        const float a = (float)pr.alpha;
        for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) {
            pr.dx[i] += a * pr.p[i];
            float ri = pr.r[i] - a * pr.q[i];
            pr.r[i] = ri;
            float zi = pr.dinv32[i] * ri;
            rz += (double)ri * (double)zi;
            rr += (double)ri * (double)ri;
        }
        block_sum2(rz, rr);
        if (threadIdx.x == 0) {
            part[2 * t] = rz;
            part[2 * t + 1] = rr;
        }
    }
}

__global__ __launch_bounds__(KB) void k_eng_elem(ENGINE_ARGS)
{
    const int nt = *ntiles;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const KTile tl = T[t];
        const int pi = tl.prob;
        const int ph = P[pi].phase;
        if (ph != PH_ITER && ph != PH_UPDATE && ph != PH_INIT && ph != PH_REINIT && ph != PH_INIT64) continue;
        const KProb pr = P[pi];
        if (ph == PH_ITER && pr.fp64) {
            const double beta = pr.beta;
            for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) pr.p64[i] = pr.dinv64[i] * pr.r64[i] + beta * pr.p64[i];
            continue;
        }
        if (ph == PH_INIT64) {
            double rz = 0.0, rr = 0.0;
            for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) {
                const double ri = pr.r64[i];
                const double zi = pr.dinv64[i] * ri;
                pr.p64[i] = zi;
                rz += ri * zi;
                rr += ri * ri;
            }
            block_sum2(rz, rr);
            if (threadIdx.x == 0) {
                part[2 * t] = rz;
                part[2 * t + 1] = rr;
            }
            continue;
        }
        if (ph == PH_ITER) {
            const float beta = (float)pr.beta;
            for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) pr.p[i] = pr.dinv32[i] * pr.r[i] + beta * pr.p[i];
            continue;
        }
        if (ph == PH_UPDATE) {
            const double inv = 1.0 / pr.scale;
            const float beta = (float)pr.beta;
            const bool pupd = pr.restart == 0;
            const bool keep = pr.discard == 0;
            for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) {
                if (keep) pr.x[i] += inv * (double)pr.dx[i];
                if (pupd) pr.p[i] = pr.dinv32[i] * pr.r[i] + beta * pr.p[i];
            }
            continue;
        }
        double rz = 0.0, rr = 0.0;
        const float ps = (float)pr.pscale;
        for (int i = tl.row0 + threadIdx.x; i < tl.row1; i += KB) {
            float ri = (float)(pr.r64[i] * pr.scale);
            float zi = pr.dinv32[i] * ri;
            pr.r[i] = ri;
            pr.dx[i] = 0.0f;
            pr.p[i] = (ph == PH_INIT) ? zi : ps * pr.p[i];
            rz += (double)ri * (double)zi;
            rr += (double)ri * (double)ri;
        }
        block_sum2(rz, rr);
        if (threadIdx.x == 0) {
            part[2 * t] = rz;
            part[2 * t + 1] = rr;
        }
    }
}

// r64 = b - A64 x on free rows of row0 .. row1 (FP64); returns the partial r.r
__device__ __forceinline__ double resid_tile(const KProb &pr, int row0, int row1, double *lds)
{
    double acc = 0.0;
    for (int r0 = row0; r0 < row1; r0 += KB) {
        const int r1 = r0 + KB < row1 ? r0 + KB : row1;
        const double s = csr_rows_lds64(pr.rowptr, pr.cols, pr.val64, pr.x, r0, r1, lds);
        const int i = r0 + (int)threadIdx.x;
        if (i < r1) {
            double ri = pr.fixed[i] ? 0.0 : pr.b[i] - s;
            pr.r64[i] = ri;
            acc += ri * ri;
        }
    }
    return acc;
}

__global__ __launch_bounds__(KB) void k_eng_resid(ENGINE_ARGS)
{
    __shared__ double lds[SPMV_CAP64];
    const int nt = *ntiles;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const KTile tl = T[t];
        const int pi = tl.prob;
        if (P[pi].phase != PH_RESID) continue;
        const KProb pr = P[pi];
        double acc = resid_tile(pr, tl.row0, tl.row1, lds), dummy = 0.0;
        block_sum2(acc, dummy);
        if (threadIdx.x == 0) part[2 * t] = acc;
    }
}

// sums the tile partials of problem s in a fixed order (valid in thread 0)
__device__ __forceinline__ void sum_parts(const KRange &rg, const double *part, double &a, double &b)
{
    a = 0.0;
    b = 0.0;
    for (int j = threadIdx.x; j < rg.nt; j += KB) {
        a += part[2 * (rg.t0 + j)];
        b += part[2 * (rg.t0 + j) + 1];
    }
    block_sum2(a, b);
}

__device__ __forceinline__ bool isfin(double v) { return v == v && !__builtin_isinf(v); }

enum { RED_A = 0, RED_B = 1, RED_C = 2 };

template <int WHICH>
__global__ __launch_bounds__(KB) void k_eng_reduce(KProb *__restrict__ P, const KRange *__restrict__ R, const double *__restrict__ part,
                                                  int nslots)
{
    for (int s = blockIdx.x; s < nslots; s += gridDim.x) {
        const KRange rg = R[s];
        if (rg.nt == 0) continue;
        const int ph = P[s].phase;
        if (WHICH == RED_A || WHICH == RED_B) {
            if (ph != PH_ITER) continue;
        } else if (ph != PH_UPDATE && ph != PH_INIT && ph != PH_REINIT && ph != PH_RESID && ph != PH_INIT64) {
            continue;
        }
        double a = 0.0, b = 0.0;
        if (!(WHICH == RED_C && ph == PH_UPDATE)) sum_parts(rg, part, a, b);
        if (threadIdx.x != 0) continue;
        KProb &w = P[s];
        if (WHICH == RED_A) {
            // a = p.q
            // Non-finite FP32 values with finite FP64 data are an overflow of
            // the inner recurrence (it can diverge at the FP64 noise floor):
            // the sweep's correction is dropped and the refresh restarts.
            const double alpha = w.rz / a;
            if (!(a > 0.0) || !isfin(a) || !isfin(alpha)) {
                w.alpha = 0.0;
                w.restart = 1;        // breakdown: refresh and restart the direction
                w.discard = !isfin(a) || !isfin(w.rz);
                w.phase = w.fp64 ? PH_RESID : PH_UPDATE;
            } else {
                w.alpha = alpha;
            }
        } else if (WHICH == RED_B) {
            // a = r.z, b = r.r of the scaled FP32 residual.  The direction
            // update p = D r + beta p happens also on a refresh step: the
            // refreshed iteration continues along the new direction.
            w.iters += 1;
            const double rn = sqrt(b);
            if (w.fp64) {
                // FP64 fallback: the recurrence residual decides when the true
                // residual is checked
                if (!isfin(a) || !isfin(b)) {
                    w.status = GRMP_ERR_NONFINITE;
                    w.phase = PH_FAIL;
                } else {
                    w.beta = a / w.rz;
                    w.rz = a;
                    if (rn <= w.rtol * w.bnorm || w.iters >= w.maxit) w.phase = PH_RESID;
                }
                continue;
            }
            if (!isfin(a) || !isfin(b)) {
                w.restart = 1;
                w.discard = 1;
                w.phase = PH_UPDATE;
            } else {
                w.beta = a / w.rz;
                w.rz = a;
                w.rmax = rn > w.rmax ? rn : w.rmax;
                // once two sweeps are complete, a sweep running past 4x their
                // mean length has lost its convergence (FP32 recurrences near
                // the FP64 floor of an ill-conditioned problem): it ends and
                // the problem finishes in FP64
                const int64_t done = w.outer - 1;
                const int64_t cur = w.iters - w.sweep0;
                const int64_t cap = done >= 2 ? max((int64_t)256, 4 * (w.sweep0 / done)) : (int64_t)1 << 62;
                if (rn > 1e6) {
                    // the scaled residual started near 1: a sweep that grew it
                    // a millionfold has diverged and is dropped
                    w.restart = 1;
                    w.discard = 1;
                    w.phase = PH_UPDATE;
                } else if (rn <= w.rtol * w.bnorm * w.scale || rn < w.delta * w.rmax || w.iters >= w.maxit) {
                    w.phase = PH_UPDATE;  // converged estimate or reliable update
                } else if (cur > cap) {
                    w.slow = 1;
                    w.phase = PH_UPDATE;
                }
            }
        } else if (ph == PH_UPDATE) {
            w.phase = PH_RESID;
        } else if (ph == PH_INIT || ph == PH_REINIT || ph == PH_INIT64) {
            w.rz = a;
            w.rmax = sqrt(b);
            w.phase = PH_ITER;
        } else {
            // ph == PH_RESID, a = ||b - A x||^2 over free rows
            if (w.outer == 0) {
                w.bb = a;
                w.bnorm = sqrt(a);
            }
            w.relres = w.bb > 0.0 ? sqrt(a / w.bb) : 0.0;
            if (!isfin(a)) {
                w.status = GRMP_ERR_NONFINITE;
                w.phase = PH_FAIL;
            } else if (a == 0.0 || w.relres <= w.rtol) {
                w.phase = PH_DONE;
            } else if (w.outer >= w.max_outer || w.iters >= w.maxit) {
                w.status = GRMP_ERR_NOCONV;
                w.phase = PH_FAIL;
            } else {
                // a refresh that did not halve the best FP64 residual restarts
                // the direction: the carried one no longer fits the refreshed
                // residual (FP32 rounding of A relative to its conditioning)
                const bool stalled = w.outer > 0 && !(w.relres <= 0.5 * w.best);
                if (w.outer == 0 || w.relres < w.best) w.best = w.relres;
                w.stall = stalled ? w.stall + 1 : 0;
                if (w.fp64 && w.stall >= 2) {
                    // FP64 CG restarts no longer halve the true residual either:
                    // x is at the FP64 floor of b - A x.  Within 100 rtol that
                    // is the attainable answer; below, the tolerance cannot be met
                    if (w.relres <= 100.0 * w.rtol) {
                        w.phase = PH_DONE;
                    } else {
                        w.status = GRMP_ERR_NOCONV;
                        w.phase = PH_FAIL;
                    }
                    continue;
                }
                // a first sweep this long marks an ill-conditioned problem, whose
                // FP32 sweeps would stall near the FP64 floor: FP64 from here
                const bool hard = w.fp64_after > 0 && w.outer == 1 && w.iters > w.fp64_after;
                if (w.fp64 || w.stall >= 2 || w.slow || hard) {
                    // two refreshes in a row without halving the residual, or a
                    // sweep past its iteration cap: the FP32 recurrences are too
                    // coarse for this problem's condition number, and FP64 CG
                    // finishes it from the current x
                    w.fp64 = 1;
                    w.phase = PH_INIT64;
                    w.outer += 1;
                    continue;
                }
                const double snew = 1.0 / sqrt(a);
                const bool fresh = w.outer == 0 || w.restart || stalled;
                w.pscale = fresh ? 1.0 : snew / w.scale;
                w.scale = snew;
                w.phase = fresh ? PH_INIT : PH_REINIT;
                w.restart = 0;
                w.discard = 0;
                w.sweep0 = w.iters;
                w.outer += 1;
            }
        }
    }
}

// --------------------------------------------------------- engine host
KEngine::KEngine(grmp_ctx *ctx, hipStream_t st, int cap_probs, int cap_tiles, int tile_rows)
    : ctx_(ctx), st_(st), cap_p_(cap_probs), cap_t_(cap_tiles), tile_rows_(tile_rows)
{
}

KEngine::~KEngine()
{
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (dP_) (void)hipFree(dP_);
    if (dT_) (void)hipFree(dT_);
    if (dR_) (void)hipFree(dR_);
    if (dpart_) (void)hipFree(dpart_);
    if (dnt_) (void)hipFree(dnt_);
    if (dbytes_) (void)hipFree(dbytes_);
    if (dbytes_blk_) (void)hipFree(dbytes_blk_);
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (hP_) (void)hipHostFree(hP_);
    if (hT_) (void)hipHostFree(hT_);
    if (hR_) (void)hipHostFree(hR_);
}

int KEngine::init()
{
    grmp_ctx *ctx = ctx_;
    GRMP_HIP(ctx, hipMalloc(&dP_, sizeof(KProb) * cap_p_));
    GRMP_HIP(ctx, hipMalloc(&dT_, sizeof(KTile) * cap_t_));
    GRMP_HIP(ctx, hipMalloc(&dR_, sizeof(KRange) * cap_p_));
    GRMP_HIP(ctx, hipMalloc(&dpart_, sizeof(double) * 2 * cap_t_));
    GRMP_HIP(ctx, hipMalloc(&dnt_, sizeof(int)));
    GRMP_HIP(ctx, hipMalloc(&dbytes_, sizeof(unsigned long long)));
    GRMP_HIP(ctx, hipMalloc(&dbytes_blk_, sizeof(unsigned long long) * ENGINE_MAX_GRID));
    GRMP_HIP(ctx, hipMemsetAsync(dbytes_, 0, sizeof(unsigned long long), st_));
    for (auto &e : ev_) GRMP_HIP(ctx, hipEventCreate(&e));
    GRMP_HIP(ctx, hipHostMalloc((void **)&hP_, sizeof(KProb) * cap_p_, hipHostMallocDefault));
    GRMP_HIP(ctx, hipHostMalloc((void **)&hT_, sizeof(KTile) * cap_t_, hipHostMallocDefault));
    GRMP_HIP(ctx, hipHostMalloc((void **)&hR_, sizeof(KRange) * cap_p_, hipHostMallocDefault));
    GRMP_HIP(ctx, hipMemsetAsync(dP_, 0, sizeof(KProb) * cap_p_, st_));
    GRMP_HIP(ctx, hipMemsetAsync(dnt_, 0, sizeof(int), st_));
    slots_.assign(cap_p_, EngineSlot());
    init_.assign(cap_p_, KProb());
    GRMP_HIP(ctx, hipStreamSynchronize(st_));
    return GRMP_OK;
}

int KEngine::add(const KProb &p, int *slot)
{
    grmp_ctx *ctx = ctx_;
    int s = -1;
    for (int k = 0; k < cap_p_; ++k)
        if (!slots_[k].used) { s = k; break; }
    GRMP_CHECK(ctx, s >= 0, GRMP_ERR_ADMISSION, "engine: all %d problem slots in use", cap_p_);
    int nt = (int)((p.n + tile_rows_ - 1) / tile_rows_);
    GRMP_CHECK(ctx, ttot_ + nt <= cap_t_, GRMP_ERR_ADMISSION, "engine: tile table full");
    KProb q = p;
    q.phase = PH_RESID;
    q.status = GRMP_OK;
    q.outer = 0;
    q.iters = 0;
    q.scale = 1.0;
    q.pscale = 1.0;
    q.rmax = 0.0;
    q.restart = 0;
    init_[s] = q;
    // the host copy lives in init_ until the stream has consumed it
    GRMP_HIP(ctx, hipMemcpyAsync(dP_ + s, &init_[s], sizeof(KProb), hipMemcpyHostToDevice, st_));
    slots_[s].used = true;
    slots_[s].ntiles = nt;
    ttot_ += nt;
    ++nact_;
    dirty_ = true;
    *slot = s;
    return GRMP_OK;
}

void KEngine::remove(int slot)
{
    if (slot < 0 || slot >= cap_p_ || !slots_[slot].used) return;
    // the slot's record goes idle on the device before its buffers can be
    // released: no kernel touches a problem whose phase is idle
    static const int32_t idle = PH_IDLE;
    (void)hipMemcpyAsync((char *)(dP_ + slot) + offsetof(KProb, phase), &idle, sizeof(int32_t), hipMemcpyHostToDevice, st_);
    slots_[slot].used = false;
    ttot_ -= slots_[slot].ntiles;
    slots_[slot].ntiles = 0;
    --nact_;
    dirty_ = true;
}

int KEngine::sync_tables()
{
    if (!dirty_) return GRMP_OK;
    grmp_ctx *ctx = ctx_;
    // the pinned staging buffers may still feed an earlier copy
    GRMP_HIP(ctx, hipStreamSynchronize(st_));
    int64_t t = 0;
    for (int s = 0; s < cap_p_; ++s) {
        if (!slots_[s].used) {
            hR_[s].t0 = 0;
            hR_[s].nt = 0;
            continue;
        }
        const int n = init_[s].n;
        hR_[s].t0 = (int32_t)t;
        hR_[s].nt = slots_[s].ntiles;
        for (int k = 0; k < slots_[s].ntiles; ++k) {
            hT_[t].prob = s;
            hT_[t].row0 = k * tile_rows_;
            hT_[t].row1 = std::min(n, (k + 1) * tile_rows_);
            hT_[t].pad = 0;
            ++t;
        }
    }
    int ntot = (int)t;
    GRMP_HIP(ctx, hipMemcpyAsync(dT_, hT_, sizeof(KTile) * (t ? t : 1), hipMemcpyHostToDevice, st_));
    GRMP_HIP(ctx, hipMemcpyAsync(dR_, hR_, sizeof(KRange) * cap_p_, hipMemcpyHostToDevice, st_));
    GRMP_HIP(ctx, hipMemcpyAsync(dnt_, &ntot, sizeof(int), hipMemcpyHostToDevice, st_));
    // ntot and the staging buffers must outlive the copies
    GRMP_HIP(ctx, hipStreamSynchronize(st_));
    dirty_ = false;
    return GRMP_OK;
}

void KEngine::launch_kernels(int grid, bool timed)
{
    dim3 g(grid), b(KB), gr(std::min(cap_p_, 1024));
    const int pr = npending_;
    if (timed) (void)hipEventRecord(ev_[2 * pr], st_);
    hipLaunchKernelGGL(k_eng_spmv, g, b, 0, st_, dP_, dT_, dR_, dpart_, dnt_, dbytes_blk_);
    if (timed) {
        (void)hipEventRecord(ev_[2 * pr + 1], st_);
        ++npending_;
        hipLaunchKernelGGL(k_eng_bytes, dim3(1), b, 0, st_, dbytes_blk_, grid, dbytes_);
    }
    hipLaunchKernelGGL(k_eng_reduce<RED_A>, gr, b, 0, st_, dP_, dR_, dpart_, cap_p_);
    hipLaunchKernelGGL(k_eng_axpy, g, b, 0, st_, dP_, dT_, dR_, dpart_, dnt_, dbytes_);
    hipLaunchKernelGGL(k_eng_reduce<RED_B>, gr, b, 0, st_, dP_, dR_, dpart_, cap_p_);
    hipLaunchKernelGGL(k_eng_elem, g, b, 0, st_, dP_, dT_, dR_, dpart_, dnt_, dbytes_);
    hipLaunchKernelGGL(k_eng_resid, g, b, 0, st_, dP_, dT_, dR_, dpart_, dnt_, dbytes_);
    hipLaunchKernelGGL(k_eng_reduce<RED_C>, gr, b, 0, st_, dP_, dR_, dpart_, cap_p_);
}

int KEngine::launch_step(int grid)
{
    launch_kernels(grid, profile_);
    GRMP_HIP(ctx_, hipGetLastError());
    ++steps_;
    if (profile_ && npending_ == EV_PAIRS) GRMP_TRY(read_events(npending_));
    return GRMP_OK;
}

int KEngine::read_events(int pairs)
{
    npending_ = 0;
    if (pairs == 0) return GRMP_OK;
    GRMP_HIP(ctx_, hipEventSynchronize(ev_[2 * pairs - 1]));
    for (int k = 0; k < pairs; ++k) {
        float ms = 0.f;
        GRMP_HIP(ctx_, hipEventElapsedTime(&ms, ev_[2 * k], ev_[2 * k + 1]));
        spmv_s_ += ms * 1e-3;
        ++spmv_n_;
    }
    return GRMP_OK;
}

int KEngine::stats(double *seconds, int64_t *launches, double *bytes, int64_t *steps)
{
    unsigned long long b = 0;
    GRMP_HIP(ctx_, hipMemcpyAsync(&b, dbytes_, sizeof(b), hipMemcpyDeviceToHost, st_));
    GRMP_HIP(ctx_, hipStreamSynchronize(st_));
    if (seconds) *seconds = spmv_s_;
    if (launches) *launches = spmv_n_;
    if (bytes) *bytes = (double)b;
    if (steps) *steps = steps_;
    return GRMP_OK;
}

#define ENGINE_GRAPH_STEPS 8

int KEngine::capture(int grid)
{
    grmp_ctx *ctx = ctx_;
    if (exec_) { (void)hipGraphExecDestroy(exec_); exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    GRMP_HIP(ctx, hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < ENGINE_GRAPH_STEPS; ++s) launch_kernels(grid, false);
    GRMP_HIP(ctx, hipStreamEndCapture(st_, &graph_));
    GRMP_HIP(ctx, hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    graph_grid_ = grid;
    return GRMP_OK;
}

int KEngine::run(int steps, bool use_graph)
{
    GRMP_TRY(sync_tables());
    if (ttot_ == 0) return GRMP_OK;
    int grid = 64;
    while (grid < ttot_ && grid < ENGINE_MAX_GRID) grid <<= 1;
    // profiling times every SpMV launch with an event pair, which a graph
    // replay cannot carry, so profiled runs launch the steps one by one
    if (use_graph && !profile_) {
        if (!exec_ || graph_grid_ != grid) GRMP_TRY(capture(grid));
        for (int s = 0; s < steps; s += ENGINE_GRAPH_STEPS) {
            GRMP_HIP(ctx_, hipGraphLaunch(exec_, st_));
            steps_ += ENGINE_GRAPH_STEPS;
        }
    } else {
        for (int s = 0; s < steps; ++s) GRMP_TRY(launch_step(grid));
        if (profile_) GRMP_TRY(read_events(npending_));
    }
    return GRMP_OK;
}

int KEngine::poll(std::vector<KProb> &out)
{
    grmp_ctx *ctx = ctx_;
    GRMP_HIP(ctx, hipMemcpyAsync(hP_, dP_, sizeof(KProb) * cap_p_, hipMemcpyDeviceToHost, st_));
    GRMP_HIP(ctx, hipStreamSynchronize(st_));
    out.assign(hP_, hP_ + cap_p_);
    return GRMP_OK;
}

// ------------------------------------------------------------ solve API
extern "C" void grmp_solve_opts_default(grmp_solve_opts *o)
{
    if (!o) return;
    o->method = GRMP_SOLVE_CG;
    o->precond = GRMP_PREC_JACOBI;
    o->precision = GRMP_PRECISION_FP64;
    o->rtol = 1e-10;
    o->maxit = 1000000;
    o->inner_rtol = 0.1;
    o->inner_maxit = 100000;
    o->check_every = 1;
    o->initial_guess = 0;
}

static int solve_engine(grmp_matrix *A, const grmp_vector *b, grmp_vector *x, const grmp_solve_opts &o, grmp_solve_info &info)
{
    grmp_ctx *ctx = A->ctx;
    const int64_t n = A->n;
    GRMP_CHECK(ctx, A->val32 && A->val64 && A->dinv32 && A->fixed && A->diagpos, GRMP_ERR_ARG,
               "FP32_IR solve needs FP64 and FP32 values assembled");
    double *r64 = nullptr;
    float *buf = nullptr;
    GRMP_TRY(dalloc(ctx, &r64, n));
    GRMP_TRY(dalloc(ctx, &buf, 4 * n));
    hipLaunchKernelGGL(k_plain_init, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, A->rowptr, A->diagpos, A->val64, A->fixed,
                       b->d, x->d, n);
    KEngine eng(ctx, ctx->stream, 1, (int)((n + ENGINE_TILE_ROWS - 1) / ENGINE_TILE_ROWS) + 1, ENGINE_TILE_ROWS);
    GRMP_TRY(eng.init());
    KProb p;
    memset(&p, 0, sizeof(p));
    p.rowptr = A->rowptr;
    p.cols = A->cols;
    p.val64 = A->val64;
    p.val32 = A->val32;
    p.dinv32 = A->dinv32;
    p.dinv64 = A->dinv64;
    p.fixed = A->fixed;
    p.b = b->d;
    p.x = x->d;
    p.r64 = r64;
    p.r = buf;
    p.p = buf + n;
    p.q = buf + 2 * n;
    p.dx = buf + 3 * n;
    p.p64 = (double *)buf;
    p.q64 = (double *)(buf + 2 * n);
    p.n = (int32_t)n;
    p.max_outer = 200;
    p.maxit = o.maxit;
    p.rtol = o.rtol;
    p.delta = o.inner_rtol;
    int slot;
    GRMP_TRY(eng.add(p, &slot));
    std::vector<KProb> st;
    int status = GRMP_OK;
    for (;;) {
        GRMP_TRY(eng.run(64, true));
        GRMP_TRY(eng.poll(st));
        const KProb &q = st[slot];
        if (q.phase == PH_DONE || q.phase == PH_FAIL) {
            status = q.phase == PH_DONE ? GRMP_OK : q.status;
            info.iterations = q.iters;
            info.outer = q.outer;
            info.relres = q.relres;
            break;
        }
    }
    info.status = status;
    dfree(ctx, r64);
    dfree(ctx, buf);
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GRMP_OK;
}

extern "C" int grmp_solve(grmp_matrix *A, const grmp_vector *b, grmp_vector *x, const grmp_solve_opts *opts,
                          grmp_solve_info *info)
{
    if (!A || !b || !x) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    grmp_solve_opts o;
    grmp_solve_opts_default(&o);
    if (opts) o = *opts;
    grmp_solve_info inf;
    memset(&inf, 0, sizeof(inf));
    GRMP_CHECK(ctx, A->have_pattern && A->val64, GRMP_ERR_ARG, "grmp_solve: matrix not assembled in FP64");
    GRMP_CHECK(ctx, b->n == A->n && x->n == A->n, GRMP_ERR_ARG, "grmp_solve: vector length != %lld", (long long)A->n);
    GRMP_CHECK(ctx, o.method == GRMP_SOLVE_CG || o.method == GRMP_SOLVE_MINRES || o.method == GRMP_SOLVE_BICGSTAB,
               GRMP_ERR_UNSUPPORTED, "grmp_solve: method %d is not built", o.method);
    GRMP_CHECK(ctx, o.method != GRMP_SOLVE_BICGSTAB || o.precision == GRMP_PRECISION_FP64, GRMP_ERR_UNSUPPORTED,
               "grmp_solve: BiCGStab runs in FP64 only");
    GRMP_CHECK(ctx, o.initial_guess == 0 || o.method == GRMP_SOLVE_BICGSTAB, GRMP_ERR_UNSUPPORTED,
               "grmp_solve: only BiCGStab starts from an initial guess");
    std::lock_guard<std::mutex> g(ctx->mu);
    double t0 = wall_seconds();
    int rc;
    if (o.method == GRMP_SOLVE_MINRES) rc = minres_solve(A, b->d, x->d, o, inf);
    else if (o.method == GRMP_SOLVE_BICGSTAB) rc = bicgstab_solve(A, b->d, x->d, o, inf);
    else if (o.precision == GRMP_PRECISION_FP32_IR) rc = solve_engine(A, b, x, o, inf);
    else rc = plain_cg_solve(A, b->d, x->d, o, inf);
    inf.seconds = wall_seconds() - t0;
    if (info) *info = inf;
    if (rc) return rc;
    if (inf.status == GRMP_ERR_NOCONV)
        return grmp_fail(ctx, GRMP_ERR_NOCONV, "grmp_solve: no convergence, relres %.3e after %lld iterations",
                         inf.relres, (long long)inf.iterations);
    if (inf.status == GRMP_ERR_NONFINITE) return grmp_fail(ctx, GRMP_ERR_NONFINITE, "grmp_solve: non-finite values in the Krylov loop");
    return inf.status;
}

// MINRES for the Stokes saddle point with the velocity's node block
// preconditioned by one multigrid V-cycle per component
extern "C" int grmp_solve_minres_mg(grmp_matrix *A, const grmp_vector *b, grmp_vector *x,
                                    const grmp_solve_opts *opts, grmp_mg *velocity, int ncomp,
                                    grmp_solve_info *info)
{
    if (!A || !b || !x || !velocity) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    grmp_solve_opts o;
    grmp_solve_opts_default(&o);
    if (opts) o = *opts;
    o.method = GRMP_SOLVE_MINRES;
    grmp_solve_info inf;
    memset(&inf, 0, sizeof(inf));
    GRMP_CHECK(ctx, A->have_pattern && A->val64, GRMP_ERR_ARG, "grmp_solve_minres_mg: matrix not assembled in FP64");
    GRMP_CHECK(ctx, b->n == A->n && x->n == A->n, GRMP_ERR_ARG, "grmp_solve_minres_mg: vector length != %lld",
               (long long)A->n);
    GRMP_CHECK(ctx, o.precision == GRMP_PRECISION_FP64 && o.initial_guess == 0, GRMP_ERR_UNSUPPORTED,
               "grmp_solve_minres_mg: FP64 from a zero start only");
    std::lock_guard<std::mutex> g(ctx->mu);
    const double t0 = wall_seconds();
    const int rc = minres_solve(A, b->d, x->d, o, inf, velocity, ncomp);
    inf.seconds = wall_seconds() - t0;
    if (info) *info = inf;
    if (rc) return rc;
    if (inf.status == GRMP_ERR_NOCONV)
        return grmp_fail(ctx, GRMP_ERR_NOCONV, "grmp_solve_minres_mg: no convergence, relres %.3e after %lld iterations",
                         inf.relres, (long long)inf.iterations);
    if (inf.status == GRMP_ERR_NONFINITE)
        return grmp_fail(ctx, GRMP_ERR_NONFINITE, "grmp_solve_minres_mg: non-finite values in the Krylov loop");
    return inf.status;
}

// BiCGStab right-preconditioned by one algebraic multigrid V-cycle per
// application (the Newton systems of nonlinear problems)
extern "C" int grmp_solve_bicgstab_amg(grmp_matrix *A, const grmp_vector *b, grmp_vector *x,
                                       const grmp_solve_opts *opts, grmp_amg *pc, grmp_solve_info *info)
{
    if (!A || !b || !x || !pc) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    grmp_solve_opts o;
    grmp_solve_opts_default(&o);
    if (opts) o = *opts;
    o.method = GRMP_SOLVE_BICGSTAB;
    grmp_solve_info inf;
    memset(&inf, 0, sizeof(inf));
    GRMP_CHECK(ctx, A->have_pattern && A->val64, GRMP_ERR_ARG, "grmp_solve_bicgstab_amg: matrix not assembled in FP64");
    GRMP_CHECK(ctx, b->n == A->n && x->n == A->n, GRMP_ERR_ARG, "grmp_solve_bicgstab_amg: vector length != %lld",
               (long long)A->n);
    GRMP_CHECK(ctx, amg_context(pc) == ctx && amg_level0_rows(pc) == A->n, GRMP_ERR_ARG,
               "grmp_solve_bicgstab_amg: the hierarchy belongs to another context or has %lld rows, not %lld",
               (long long)amg_level0_rows(pc), (long long)A->n);
    GRMP_CHECK(ctx, o.precision == GRMP_PRECISION_FP64, GRMP_ERR_UNSUPPORTED, "grmp_solve_bicgstab_amg: FP64 only");
    std::lock_guard<std::mutex> g(ctx->mu);
    const double t0 = wall_seconds();
    const int rc = bicgstab_solve(A, b->d, x->d, o, inf, pc);
    inf.seconds = wall_seconds() - t0;
    if (info) *info = inf;
    if (rc) return rc;
    if (inf.status == GRMP_ERR_NOCONV)
        return grmp_fail(ctx, GRMP_ERR_NOCONV, "grmp_solve_bicgstab_amg: no convergence, relres %.3e after %lld iterations",
                         inf.relres, (long long)inf.iterations);
    if (inf.status == GRMP_ERR_NONFINITE)
        return grmp_fail(ctx, GRMP_ERR_NONFINITE, "grmp_solve_bicgstab_amg: non-finite values in the Krylov loop");
    return inf.status;
}

// ------------------------------------------------------- CSC plug-in solve
extern "C" int grmp_solve_csc(grmp_ctx *ctx, int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval,
                              const double *b, double *x, int64_t nfixed, const int64_t *fixed, int precision, double rtol)
{
    GRMP_CHECK(ctx, ctx && n > 0 && n < (1ll << 31) && colptr && rowval && nzval && b && x, GRMP_ERR_ARG,
               "grmp_solve_csc: bad arguments");
    const int64_t nnz = colptr[n] - 1;
    GRMP_CHECK(ctx, nnz >= 0 && nnz < (1ll << 31), GRMP_ERR_ARG, "grmp_solve_csc: nnz out of range");
    // CSC (columns) -> CSR (rows), 0-based, rows sorted by column
    std::vector<int32_t> rp(n + 1, 0), cl(nnz);
    std::vector<double> v64(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
        GRMP_CHECK(ctx, rowval[k] >= 1 && rowval[k] <= n, GRMP_ERR_ARG, "grmp_solve_csc: row index out of range");
        rp[rowval[k]]++;
    }
    for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<int32_t> cur(rp.begin(), rp.end() - 1);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) {
            int32_t p = cur[rowval[k] - 1]++;
            cl[p] = (int32_t)j;
            v64[p] = nzval[k];
        }
    std::vector<uint8_t> fx(n, 0), dpos(n, 255);
    for (int64_t k = 0; k < nfixed; ++k) {
        GRMP_CHECK(ctx, fixed[k] >= 1 && fixed[k] <= n, GRMP_ERR_ARG, "grmp_solve_csc: fixed dof out of range");
        fx[fixed[k] - 1] = 1;
    }
    std::vector<double> di64(n, 0.0);
    std::vector<float> di32(n, 0.f), v32(nnz);
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            if (cl[k] == i) {
                d = v64[k];
                if (k - rp[i] < 255) dpos[i] = (uint8_t)(k - rp[i]);
            }
            v32[k] = fx[i] ? 0.f : (float)v64[k];
        }
        GRMP_CHECK(ctx, fx[i] == 0 || dpos[i] != 255, GRMP_ERR_ARG, "grmp_solve_csc: fixed dof without a diagonal entry");
        di64[i] = (fx[i] || d == 0.0) ? 0.0 : 1.0 / d;
        di32[i] = (float)di64[i];
    }
    grmp_matrix A;
    A.ctx = ctx;
    A.nblocks = 0;
    A.n = n;
    A.nnz = nnz;
    A.have_pattern = true;
    int rc = GRMP_OK;
    grmp_vector bv{ctx, n, nullptr}, xv{ctx, n, nullptr};
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        GRMP_TRY(dalloc(ctx, &A.rowptr, n + 1));
        GRMP_TRY(dalloc(ctx, &A.cols, nnz));
        GRMP_TRY(dalloc(ctx, &A.val64, nnz));
        GRMP_TRY(dalloc(ctx, &A.val32, nnz));
        GRMP_TRY(dalloc(ctx, &A.dinv64, n));
        GRMP_TRY(dalloc(ctx, &A.dinv32, n));
        GRMP_TRY(dalloc(ctx, &A.diagpos, n));
        GRMP_TRY(dalloc(ctx, &A.fixed, n));
        GRMP_TRY(dalloc(ctx, &bv.d, n));
        GRMP_TRY(dalloc(ctx, &xv.d, n));
        hipStream_t st = ctx->stream;
        GRMP_HIP(ctx, hipMemcpyAsync(A.rowptr, rp.data(), 4 * (n + 1), hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.cols, cl.data(), 4 * nnz, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.val64, v64.data(), 8 * nnz, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.val32, v32.data(), 4 * nnz, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.dinv64, di64.data(), 8 * n, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.dinv32, di32.data(), 4 * n, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.diagpos, dpos.data(), n, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(A.fixed, fx.data(), n, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipMemcpyAsync(bv.d, b, 8 * n, hipMemcpyHostToDevice, st));
        GRMP_HIP(ctx, hipStreamSynchronize(st));
    }
    grmp_solve_opts o;
    grmp_solve_opts_default(&o);
    for (int64_t i = 0; i < n; ++i)
        if (!fx[i] && di64[i] == 0.0) {
            o.method = GRMP_SOLVE_MINRES;  // constraint rows: the system is a saddle point
            break;
        }
    o.precision = precision == GRMP_PRECISION_FP32_IR ? GRMP_PRECISION_FP32_IR : GRMP_PRECISION_FP64;
    o.rtol = rtol > 0 ? rtol : 1e-10;
    grmp_solve_info info;
    rc = grmp_solve(&A, &bv, &xv, &o, &info);
    if (rc == GRMP_OK || rc == GRMP_ERR_NOCONV) {
        int rc2 = grmp_vector_download(&xv, x);
        if (rc == GRMP_OK) rc = rc2;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    dfree(ctx, A.rowptr);
    dfree(ctx, A.cols);
    dfree(ctx, A.val64);
    dfree(ctx, A.val32);
    dfree(ctx, A.dinv64);
    dfree(ctx, A.dinv32);
    dfree(ctx, A.diagpos);
    dfree(ctx, A.fixed);
    dfree(ctx, bv.d);
    dfree(ctx, xv.d);
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
}
