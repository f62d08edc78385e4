// Preconditioned MINRES for the Stokes saddle point on the device.
//
// Replaces the direct factorization of solve_direct! (src/solvers.jl:622-686)
// for systems with constraint rows, as assembled by
// IncompressibleNavierStokesProblem(2; nonlinear=false)
// (src/pdeprototypes.jl:17-55): [[nu K, -B^T], [+B, 0]] plus the penalty on
// the no-slip and pinned pressure dofs.  The LagrangeMultiplier's transposed
// copy has the opposite sign (src/assemblypatterns/bilinearform.jl:358-363),
// so the rows whose diagonal is zero (the constraint rows) are negated on the
// fly, which makes the operator symmetric without touching the assembled
// values.
//
// Preconditioner (symmetric positive definite, diagonal):
//   rows with a diagonal:   1 / |A_ii|
//   constraint rows:        1 / sum_j A_ij^2 / A_jj  (the diagonal of
//                            B diag(K)^-1 B^T, a Schur-complement estimate)
// stored as one signed array: its sign is the row's sign flip.
// Block variant (grmp_solve_minres_mg): the node rows of each velocity
// component -- the P1 part of H1BR{2}, on which the viscous block is nu
// times the P1 Laplacian -- get one multigrid V-cycle of that Laplacian in
// place of the diagonal; the face-bubble rows keep 1 / |A_ii| and the
// constraint rows the Schur estimate.  Block diagonal and symmetric positive
// definite, so MINRES's recurrence is unchanged.
//
// The recurrence is the preconditioned Lanczos / Givens form of MINRES
// (Paige & Saunders 1975, as written in Elman, Silvester & Wathen, "Finite
// Elements and Fast Iterative Solvers", preconditioned MINRES).  Its scalars
// live on the device: a one-block kernel reduces each inner product in a
// fixed order and advances the Givens rotation, so the host only polls a
// status word every few iterations and results do not depend on the polling.
// Fixed dofs are eliminated as in the CG solvers (x_d = b_d / A_dd).
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>

#include "grmp_internal.h"
#include "grmp_sparse.h"

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#define MR_GRID 1024

struct MrScal {
    double g, gp;          // gamma_j, gamma_{j-1}
    double c, cp, s, sp;   // Givens c_j, c_{j-1}, s_j, s_{j-1}
    double eta, g1, tol;   // eta, gamma_1, stop when |eta| <= tol
    double delta;
    double u_ig, u_a1, u_a2, u_a3, u_ceta;  // deferred w/x update
    int64_t it, maxit;
    int status;            // 0 running, 1 final update pending, 2 finished
    int pending;           // a deferred w/x update exists
    int fail;              // GRMP_OK / GRMP_ERR_NOCONV / GRMP_ERR_NONFINITE
    int pad_;
};

__device__ __forceinline__ bool isfin_d(double v) { return v == v && fabs(v) <= 1.7976931348623157e308; }

// sum of the first n partials in a fixed order; valid in thread 0
__device__ __forceinline__ double fixed_sum(const double *__restrict__ part, int n)
{
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < n; k += KB) a += part[k];
    block_sum2(a, b);
    return a;
}

// ------------------------------------------------- preconditioner diagonal
// dsg[i] = 0 on fixed rows; +1/|A_ii| where the row has a nonzero diagonal;
// -1 / sum_j A_ij^2 dinv_j on the other (constraint) rows.
__global__ void k_mr_diag(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci, const double *__restrict__ v,
                          const uint8_t *__restrict__ fixed, const double *__restrict__ dinv, double *__restrict__ dsg,
                          float *__restrict__ dsg32, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    if (!fixed[i]) {
        if (dinv[i] != 0.0) {
            d = fabs(dinv[i]);
        } else {
            double s = 0.0;
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * v[k] * fabs(dinv[ci[k]]);
            d = s > 0.0 ? -1.0 / s : 0.0;
        }
    }
    dsg[i] = d;
    if (dsg32) dsg32[i] = (float)d;
}

// ------------------------------------------------------------ init
// v1 = S (rhs - A x) on free rows, z1 = |d| v1, partial z1.v1; with
// A == nullptr the residual is rhs itself (inner solves start from 0).
template <class T, class V, class X, class R>
__global__ __launch_bounds__(KB) void k_mr_init(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                const V *__restrict__ a, const uint8_t *__restrict__ fixed,
                                                const T *__restrict__ dsg, const R *__restrict__ rhs,
                                                const X *__restrict__ x, T *__restrict__ v1, T *__restrict__ z1,
                                                T *__restrict__ v0, T *__restrict__ w0, T *__restrict__ w1, int64_t n,
                                                double *__restrict__ part)
{
    __shared__ double lds[SPMV_CAP64];
    double acc = 0.0, dummy = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * KB; r0 < n; r0 += (int64_t)gridDim.x * KB) {
        const int64_t r1 = r0 + KB < n ? r0 + KB : n;
        const double s = a ? csr_rows_lds64(rp, ci, a, x, r0, r1, lds) : 0.0;
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) {
            const double d = (double)dsg[i];
            double r = fixed[i] ? 0.0 : ((double)rhs[i] - s);
            if (d < 0.0) r = -r;
            const T vt = (T)r;
            const T zt = (T)(fabs(d) * (double)vt);
            v1[i] = vt;
            z1[i] = zt;
            v0[i] = (T)0;
            w0[i] = (T)0;
            w1[i] = (T)0;
            acc += (double)zt * (double)vt;
        }
    }
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(KB) void k_mr_start(MrScal *__restrict__ S, const double *__restrict__ part, int np,
                                                 double rtol, int64_t maxit)
{
    const double gg = fixed_sum(part, np);
    if (threadIdx.x != 0) return;
    MrScal s;
    memset(&s, 0, sizeof(s));
    const double g = sqrt(fmax(gg, 0.0));
    s.g = g;
    s.gp = 1.0;
    s.c = s.cp = 1.0;
    s.s = s.sp = 0.0;
    s.eta = g;
    s.g1 = g;
    s.tol = rtol * g;
    s.maxit = maxit;
    s.status = (g == 0.0) ? 2 : 0;
    s.fail = isfin_d(gg) ? GRMP_OK : GRMP_ERR_NONFINITE;
    if (s.fail) s.status = 2;
    *S = s;
}

// ------------------------------------------------------------ iteration
// K1: deferred w/x update of the previous iteration, then
//     q = S A z_j / gamma_j (0 on fixed rows), partial q.z_j / gamma_j
template <class T, class V, class X>
__global__ __launch_bounds__(KB) void k_mr_apply(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                 const V *__restrict__ a, const uint8_t *__restrict__ fixed,
                                                 const T *__restrict__ dsg, const T *__restrict__ z,
                                                 const T *__restrict__ zprev, T *__restrict__ wnew,
                                                 const T *__restrict__ wm2, const T *__restrict__ wm1,
                                                 X *__restrict__ x, T *__restrict__ q, int64_t n,
                                                 const MrScal *__restrict__ S, double *__restrict__ part)
{
    __shared__ T lds[SPMV_CAP];
    const int status = S->status;
    if (status == 2) return;
    const bool pend = S->pending != 0;
    const double uig = S->u_ig, ua1 = S->u_a1, ua2 = S->u_a2, ua3 = S->u_a3, uce = S->u_ceta;
    const double ig = 1.0 / S->g;
    double acc = 0.0, dummy = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * KB; r0 < n; r0 += (int64_t)gridDim.x * KB) {
        const int64_t r1 = r0 + KB < n ? r0 + KB : n;
        const int64_t i = r0 + threadIdx.x;
        if (pend && i < r1) {
            const double wn = ((double)zprev[i] * uig - ua3 * (double)wm2[i] - ua2 * (double)wm1[i]) / ua1;
            wnew[i] = (T)wn;
            x[i] = (X)((double)x[i] + uce * wn);
        }
        if (status == 0) {
            const T s = csr_rows_lds<T>(rp, ci, a, z, r0, r1, lds);
            if (i < r1) {
                const double d = (double)dsg[i];
                double qi = fixed[i] ? 0.0 : (double)s * ig;
                if (d < 0.0) qi = -qi;
                const T qt = (T)qi;
                q[i] = qt;
                acc += (double)qt * (double)z[i] * ig;
            }
        }
    }
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// K2: delta_j; a finished final update ends the solve
__global__ __launch_bounds__(KB) void k_mr_delta(MrScal *__restrict__ S, const double *__restrict__ part, int np)
{
    const int status = S->status;
    if (status == 2) return;
    const double d = fixed_sum(part, np);
    if (threadIdx.x != 0) return;
    if (status == 1) {
        S->status = 2;
        return;
    }
    S->delta = d;
}

// K3: v_{j+1} = q - (delta/gamma_j) v_j - (gamma_j/gamma_{j-1}) v_{j-1}
//     (written over v_{j-1}); z_{j+1} = |d| v_{j+1}; partial z.v
template <class T>
__global__ __launch_bounds__(KB) void k_mr_lanczos(const T *__restrict__ q, const T *__restrict__ vj, T *__restrict__ vm1,
                                                   const T *__restrict__ dsg, T *__restrict__ znew, int64_t n,
                                                   const MrScal *__restrict__ S, double *__restrict__ part)
{
    if (S->status != 0) return;
    const double c1 = S->delta / S->g, c2 = S->g / S->gp;
    double acc = 0.0, dummy = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB) {
        const double vn = (double)q[i] - c1 * (double)vj[i] - c2 * (double)vm1[i];
        const T vt = (T)vn;
        const T zt = (T)(fabs((double)dsg[i]) * (double)vt);
        vm1[i] = vt;
        znew[i] = zt;
        acc += (double)zt * (double)vt;
    }
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// K4: gamma_{j+1}, the Givens rotation and the coefficients of the deferred
//     w_{j+1} = (z_j/gamma_j - a3 w_{j-1} - a2 w_j)/a1, x += c_{j+1} eta w_{j+1}
__global__ __launch_bounds__(KB) void k_mr_givens(MrScal *__restrict__ S, const double *__restrict__ part, int np)
{
    if (S->status != 0) return;
    const double gg = fixed_sum(part, np);
    if (threadIdx.x != 0) return;
    MrScal s = *S;
    const double gn = sqrt(fmax(gg, 0.0));
    const double a0 = s.c * s.delta - s.cp * s.s * s.g;
    const double a1 = sqrt(a0 * a0 + gn * gn);
    const double a2 = s.s * s.delta + s.cp * s.c * s.g;
    const double a3 = s.sp * s.g;
    if (!isfin_d(gg) || !isfin_d(a1) || !isfin_d(s.delta) || a1 == 0.0) {
        s.fail = GRMP_ERR_NONFINITE;
        s.status = 2;
        *S = s;
        return;
    }
    const double cn = a0 / a1, sn = gn / a1;
    s.u_ig = 1.0 / s.g;
    s.u_a1 = a1;
    s.u_a2 = a2;
    s.u_a3 = a3;
    s.u_ceta = cn * s.eta;
    s.pending = 1;
    s.eta = -sn * s.eta;
    s.gp = s.g;
    s.g = gn;
    s.cp = s.c;
    s.c = cn;
    s.sp = s.s;
    s.s = sn;
    s.it += 1;
    if (fabs(s.eta) <= s.tol || gn == 0.0) {
        s.status = 1;
    } else if (s.it >= s.maxit) {
        s.status = 1;
        s.fail = GRMP_ERR_NOCONV;
    }
    *S = s;
}

// partial z.v after the block preconditioner has rewritten z
__global__ __launch_bounds__(KB) void k_mr_dot(const double *__restrict__ z, const double *__restrict__ v, int64_t n,
                                              double *__restrict__ part)
{
    double acc = 0.0, dummy = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB) acc += z[i] * v[i];
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// ------------------------------------------------------- FP64 refinement
// r = S (b - A x) on free rows (FP64), partial r.|d| r
__global__ __launch_bounds__(KB) void k_mr_resid64(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                   const double *__restrict__ a, const uint8_t *__restrict__ fixed,
                                                   const double *__restrict__ dsg, const double *__restrict__ b,
                                                   const double *__restrict__ x, double *__restrict__ r, int64_t n,
                                                   double *__restrict__ part)
{
    __shared__ double lds[SPMV_CAP64];
    double acc = 0.0, dummy = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * KB; r0 < n; r0 += (int64_t)gridDim.x * KB) {
        const int64_t r1 = r0 + KB < n ? r0 + KB : n;
        const double s = csr_rows_lds64(rp, ci, a, x, r0, r1, lds);
        const int64_t i = r0 + threadIdx.x;
        if (i < r1) {
            const double d = dsg[i];
            double ri = fixed[i] ? 0.0 : b[i] - s;
            r[i] = ri;
            acc += ri * fabs(d) * ri;
        }
    }
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void k_mr_addx(double *__restrict__ x, float *__restrict__ dx, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += (double)dx[i];
    dx[i] = 0.f;
}

__global__ void k_mr_init_x(const int32_t *rp, const uint8_t *dpos, const double *val, const uint8_t *fixed,
                            const double *b, double *x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fixed[i]) {
        const double d = dpos[i] != 255 ? val[rp[i] + dpos[i]] : 1.0;
        x[i] = b[i] / d;
    } else {
        x[i] = 0.0;
    }
}

// ------------------------------------------------------------ host side
namespace {

// z_k and v_k live in slot k % 2 (z_{j+1} and v_{j+1} are written element by
// element over z_{j-1} and v_{j-1}); w_k lives in slot k % 3.
template <class T>
struct MrBufs {
    T *Z[2], *V[2], *W[3], *q;
};

template <class T>
int mr_alloc(grmp_ctx *ctx, MrBufs<T> &B, T **block, int64_t n)
{
    GRMP_TRY(dalloc(ctx, block, 8 * n));
    T *p = *block;
    for (int k = 0; k < 2; ++k) B.Z[k] = p + k * n;
    for (int k = 0; k < 2; ++k) B.V[k] = p + (2 + k) * n;
    for (int k = 0; k < 3; ++k) B.W[k] = p + (4 + k) * n;
    B.q = p + 7 * n;
    return GRMP_OK;
}

struct MrHost {
    grmp_ctx *ctx;
    hipStream_t st;
    MrScal *dS;
    MrScal *hS;
    double *part;
    int grid, gride;
};

// the multigrid block of the block preconditioner: components 0..ncomp-1
// occupy rows [k nv, (k+1) nv)
struct MrPrec {
    grmp_mg *mg = nullptr;
    int ncomp = 0;
    int64_t nv = 0;
};

// z = P v on the multigrid rows (the diagonal part is already in z), then the
// np partials of z.v
static int mr_precondition(MrHost &H, const MrPrec &pc, const double *v, double *z, int64_t n, int np)
{
    if (pc.ncomp == 2)  // both components' V-cycles at once, on two copies of the hierarchy
        GRMP_TRY(mg_precond_apply_pair(pc.mg, v, z, v + pc.nv, z + pc.nv, H.st));
    else
        for (int k = 0; k < pc.ncomp; ++k) GRMP_TRY(mg_precond_apply(pc.mg, v + k * pc.nv, z + k * pc.nv, H.st));
    hipLaunchKernelGGL(k_mr_dot, dim3(np), dim3(KB), 0, H.st, z, v, n, H.part);
    return GRMP_OK;
}


// one MINRES run on S A y = rhs from y = y0 (the residual of y0 against rhs
// is formed when Ainit != nullptr, else y0 must be 0); returns via hS
template <class T, class V, class X, class R>
int mr_run(MrHost &H, const grmp_matrix *A, const V *vals, const T *dsg, const R *rhs, X *y, bool y_nonzero,
           MrBufs<T> &B, double rtol, int64_t maxit, int check_every, int64_t *iters_out, const MrPrec *pc = nullptr)
{
    const int64_t n = A->n;
    const V *ainit = y_nonzero ? vals : nullptr;
    hipLaunchKernelGGL((k_mr_init<T, V, X, R>), dim3(H.grid), dim3(KB), 0, H.st, A->rowptr, A->cols, ainit, A->fixed,
                       dsg, rhs, y, B.V[1], B.Z[1], B.V[0], B.W[0], B.W[1], n, H.part);
    if constexpr (std::is_same<T, double>::value) {
        if (pc) GRMP_TRY(mr_precondition(H, *pc, B.V[1], B.Z[1], n, H.grid));
    }
    hipLaunchKernelGGL(k_mr_start, dim3(1), dim3(KB), 0, H.st, H.dS, H.part, H.grid, rtol, maxit);
    GRMP_HIP(H.ctx, hipGetLastError());
    const int every = std::max(1, check_every);
    for (int64_t j = 1;; ++j) {
        T *z = B.Z[j & 1], *zprev = B.Z[(j + 1) & 1];
        T *wnew = B.W[j % 3], *wm2 = B.W[(j + 1) % 3], *wm1 = B.W[(j + 2) % 3];
        T *vj = B.V[j & 1], *vm1 = B.V[(j + 1) & 1];
        T *znew = zprev;  // z_{j+1} replaces z_{j-1}
        hipLaunchKernelGGL((k_mr_apply<T, V, X>), dim3(H.grid), dim3(KB), 0, H.st, A->rowptr, A->cols, vals, A->fixed,
                           dsg, z, zprev, wnew, wm2, wm1, y, B.q, n, H.dS, H.part);
        hipLaunchKernelGGL(k_mr_delta, dim3(1), dim3(KB), 0, H.st, H.dS, H.part, H.grid);
        hipLaunchKernelGGL(k_mr_lanczos<T>, dim3(H.gride), dim3(KB), 0, H.st, B.q, vj, vm1, dsg, znew, n, H.dS, H.part);
        if constexpr (std::is_same<T, double>::value) {
            // after the last iteration z is not read again; the extra V-cycles until the next poll are wasted work
            if (pc) GRMP_TRY(mr_precondition(H, *pc, vm1, znew, n, H.gride));
        }
        hipLaunchKernelGGL(k_mr_givens, dim3(1), dim3(KB), 0, H.st, H.dS, H.part, H.gride);
        if (j % every == 0) {
            GRMP_HIP(H.ctx, hipGetLastError());
            GRMP_HIP(H.ctx, hipMemcpyAsync(H.hS, H.dS, sizeof(MrScal), hipMemcpyDeviceToHost, H.st));
            GRMP_HIP(H.ctx, hipStreamSynchronize(H.st));
            if (H.hS->status == 2) break;
        }
    }
    *iters_out = H.hS->it;
    return GRMP_OK;
}

}  // namespace

int minres_solve(grmp_matrix *A, const double *bd, double *xd, const grmp_solve_opts &o, grmp_solve_info &info,
                 grmp_mg *vel_mg, int vel_ncomp)
{
    grmp_ctx *ctx = A->ctx;
    const int64_t n = A->n;
    const bool ir = o.precision == GRMP_PRECISION_FP32_IR;
    GRMP_CHECK(ctx, A->rowptr && A->cols && A->val64 && A->dinv64 && A->fixed && A->diagpos && bd && xd, GRMP_ERR_ARG,
               "MINRES: matrix not assembled in FP64");
    GRMP_CHECK(ctx, !ir || A->val32, GRMP_ERR_ARG, "MINRES FP32_IR: FP32 values not assembled");
    GRMP_CHECK(ctx, !vel_mg || !ir, GRMP_ERR_UNSUPPORTED, "MINRES with a multigrid block runs in FP64");
    MrPrec pc;
    if (vel_mg) {
        pc.mg = vel_mg;
        pc.ncomp = vel_ncomp;
        pc.nv = mg_level0_rows(vel_mg);
        GRMP_CHECK(ctx, vel_ncomp >= 1 && vel_ncomp <= 3 && pc.nv * vel_ncomp <= n, GRMP_ERR_ARG,
                   "MINRES: %d velocity components of %lld rows exceed the system's %lld", vel_ncomp,
                   (long long)pc.nv, (long long)n);
        if (vel_ncomp == 2) GRMP_TRY(mg_pair_prepare(vel_mg));
    }
    MrHost H;
    H.ctx = ctx;
    H.st = ctx->stream;
    H.grid = (int)std::min<int64_t>(MR_GRID, std::max<int64_t>(1, (n + KB - 1) / KB));
    H.gride = (int)std::min<int64_t>(MR_GRID, std::max<int64_t>(1, (n + KB - 1) / KB));
    H.dS = nullptr;
    H.hS = nullptr;
    H.part = nullptr;
    double *dsg = nullptr, *r64 = nullptr, *hres = nullptr;
    float *dsg32 = nullptr, *dx = nullptr;
    double *blk64 = nullptr;
    float *blk32 = nullptr;
    int rc = GRMP_OK;
    int64_t iters = 0, outer = 0;
    double relres = 0.0;
    int status = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &H.dS, 1))) break;
        if ((rc = dalloc(ctx, &H.part, MR_GRID))) break;
        if ((rc = dalloc(ctx, &dsg, n))) break;
        if (hipHostMalloc((void **)&H.hS, sizeof(MrScal), hipHostMallocDefault) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "MINRES: pinned status allocation failed");
            break;
        }
        if (ir) {
            if ((rc = dalloc(ctx, &dsg32, n))) break;
            if ((rc = dalloc(ctx, &dx, n))) break;
            if ((rc = dalloc(ctx, &r64, n))) break;
            if (hipHostMalloc((void **)&hres, sizeof(double) * MR_GRID, hipHostMallocDefault) != hipSuccess) {
                rc = grmp_fail(ctx, GRMP_ERR_HIP, "MINRES: pinned allocation failed");
                break;
            }
        }
        hipLaunchKernelGGL(k_mr_diag, dim3(nblk(n, 256)), dim3(256), 0, H.st, A->rowptr, A->cols, A->val64, A->fixed,
                           A->dinv64, dsg, dsg32, n);
        hipLaunchKernelGGL(k_mr_init_x, dim3(nblk(n, 256)), dim3(256), 0, H.st, A->rowptr, A->diagpos, A->val64, A->fixed,
                           bd, xd, n);
        if (!ir) {
            MrBufs<double> B;
            if ((rc = mr_alloc(ctx, B, &blk64, n))) break;
            if ((rc = mr_run<double, double, double, double>(H, A, A->val64, dsg, bd, xd, true, B, o.rtol, o.maxit,
                                                             vel_mg ? std::min(o.check_every, 4) : o.check_every,
                                                             &iters, vel_mg ? &pc : nullptr)))
                break;
            status = H.hS->fail;
            relres = H.hS->g1 > 0 ? fabs(H.hS->eta) / H.hS->g1 : 0.0;
            outer = 1;
        } else {
            MrBufs<float> B;
            if ((rc = mr_alloc(ctx, B, &blk32, n))) break;
            GRMP_HIP(ctx, hipMemsetAsync(dx, 0, sizeof(float) * n, H.st));
            const double delta = (o.inner_rtol > 0 && o.inner_rtol < 1) ? o.inner_rtol : 1e-3;
            double g0 = -1.0;
            for (;;) {
                // FP64 residual in the preconditioner's norm
                hipLaunchKernelGGL(k_mr_resid64, dim3(H.grid), dim3(KB), 0, H.st, A->rowptr, A->cols, A->val64, A->fixed,
                                   dsg, bd, xd, r64, n, H.part);
                GRMP_HIP(ctx, hipMemcpyAsync(hres, H.part, sizeof(double) * H.grid, hipMemcpyDeviceToHost, H.st));
                GRMP_HIP(ctx, hipStreamSynchronize(H.st));
                double rr = 0.0;
                for (int k = 0; k < H.grid; ++k) rr += hres[k];
                if (!(rr == rr) || std::isinf(rr)) {
                    status = GRMP_ERR_NONFINITE;
                    break;
                }
                const double rn = std::sqrt(rr);
                if (g0 < 0) g0 = rn;
                relres = g0 > 0 ? rn / g0 : 0.0;
                if (rn <= o.rtol * g0 || rn == 0.0) break;
                if (iters >= o.maxit || outer >= 200) {
                    status = GRMP_ERR_NOCONV;
                    break;
                }
                int64_t it = 0;
                const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(o.inner_maxit, o.maxit - iters));
                if ((rc = mr_run<float, float, float, double>(H, A, A->val32, dsg32, r64, dx, false, B, delta, cap,
                                                              o.check_every, &it)))
                    break;
                iters += it;
                ++outer;
                if (H.hS->fail == GRMP_ERR_NONFINITE) {
                    status = GRMP_ERR_NONFINITE;
                    break;
                }
                hipLaunchKernelGGL(k_mr_addx, dim3(nblk(n, 256)), dim3(256), 0, H.st, xd, dx, n);
            }
        }
    } while (0);
    (void)hipStreamSynchronize(ctx->stream);
    info.iterations = iters;
    info.outer = outer;
    info.relres = relres;
    info.status = rc ? rc : status;
    if (H.hS) (void)hipHostFree(H.hS);
    if (hres) (void)hipHostFree(hres);
    dfree(ctx, H.dS);
    dfree(ctx, H.part);
    dfree(ctx, dsg);
    dfree(ctx, dsg32);
    dfree(ctx, dx);
    dfree(ctx, r64);
    dfree(ctx, blk64);
    dfree(ctx, blk32);
    return rc;
}
