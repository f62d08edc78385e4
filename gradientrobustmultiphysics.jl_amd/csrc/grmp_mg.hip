// Multigrid-preconditioned CG for the served unit-square families (Poisson
// P1 / P2, elasticity P1 with a per-cell coefficient, homogeneous Dirichlet
// data on the whole boundary, penalized fixed rows as in the reference's
// solve_direct!, src/solvers.jl:647-652).
//
// The reference solves these systems with a sparse factorization or an
// unpreconditioned Krylov method (src/solvers.jl:38-59, 622-660); the
// solution it defines is the FP64 solution of the penalized system, which this
// solver reaches to the requested relative residual.  The preconditioner is
// this library's own: one V-cycle of a geometric hierarchy
//   level 0      the request's own matrix (P1, P2 or P1 x 2 on N x N),
//   [level 1     P1 on the same mesh, when level 0 is P2],
//   then         P1 on ceil(N_l / 2) x ceil(N_l / 2), down to N_l <= 4,
// each coarse matrix assembled by the same row-owner kernels from a
// coefficient coarsened by the geometric mean of the fine cells whose
// centroid lies in the coarse cell.  Transfers are linear interpolation on the
// coarse triangulation (nested or not) and its exact transpose, both written
// as gathers so that every sum has a fixed order; fixed dofs receive no
// correction and carry no residual.  Smoothing is Chebyshev of degree 2 on
// D^-1 A over [lmax/8, lmax] with lmax the Gershgorin bound of D^-1 A; the
// coarsest level is solved exactly with its dense inverse.  Pre- and
// post-smoothing are the same polynomial, so the V-cycle is symmetric and CG
// applies.  All scalars live on the device; the host only polls a status
// copy.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "grmp_internal.h"
#include "grmp_sparse.h"
#include "grmp_square.h"

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Start / stop events of a profiled launch come from a pool that collecting
// refills, so hipEventCreate runs only while the pool grows; every creation
// is checked, and a launch whose pair cannot be created runs unprofiled.
static bool ev_take(std::vector<hipEvent_t> &pool, hipEvent_t &e0, hipEvent_t &e1)
{
    e0 = e1 = nullptr;
    while (pool.size() < 2) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess || e == nullptr) return false;
        pool.push_back(e);
    }
    e1 = pool.back();
    pool.pop_back();
    e0 = pool.back();
    pool.pop_back();
    return true;
}

#define MG_MAX_LEVELS 16
#define MG_DENSE_MAX 64  // dofs of the coarsest level (P1 x 2 on 4 x 4: 50)

// Chebyshev degree 2 on [lmax/8, lmax]: theta = 9/16 lmax, delta = 7/16 lmax,
// rho0 = 7/9, rho1 = 63/113.
//   x1 = x0 + d0,  d0 = D^-1 r0 / theta
//   x2 = x1 + C1 d0 + C2 D^-1 (b - A x1),  C1 = rho1 rho0, C2 = 2 rho1 / delta
#define CHEB_ITHETA(l) (16.0 / (9.0 * (l)))
#define CHEB_C1 (49.0 / 113.0)
#define CHEB_C2(l) (288.0 / (113.0 * (l)))

// done: 0 iterating, 2 the updated residual met rtol (the true residual is
// checked next), 1 finished (status says how)
struct MgScal {
    double rz, pq, rr, rr0, alpha, beta, relres, rtol;
    int it, done, status, maxit;
    double best;      // smallest true relative residual seen at a check
    int restarts, pad;
};

enum MgRed { RED_INIT = 0, RED_ALPHA = 1, RED_CONV = 2, RED_BETA = 3, RED_VERIFY = 4, RED_RESTART = 5 };
// the scalar step of W on the reduced values a (and b for RED_INIT); one thread
template <int W>
__device__ __forceinline__ void mg_scalar(MgScal *sc, double a, double b)
{
    if (W == RED_INIT) {  // a = r.z, b = r.r of the start residual
        sc->rz = a;
        sc->rr0 = b;
        sc->rr = b;
        sc->beta = 0.0;
        sc->relres = b > 0.0 ? 1.0 : 0.0;
        if (!(b > 0.0)) sc->done = 1;  // zero right-hand side: x = 0
        else if (!(a > 0.0) || !isfinite(a)) { sc->status = GRMP_ERR_NONFINITE; sc->done = 1; }
    } else if (W == RED_ALPHA) {
        sc->pq = a;
        if (!(a > 0.0) || !isfinite(a)) { sc->status = GRMP_ERR_NONFINITE; sc->done = 1; return; }
        sc->alpha = sc->rz / a;
    } else if (W == RED_CONV) {
        sc->rr = a;
        sc->it += 1;
        sc->relres = sqrt(a / sc->rr0);
        if (!isfinite(a)) { sc->status = GRMP_ERR_NONFINITE; sc->done = 1; }
        else if (sc->relres <= sc->rtol) sc->done = 2;
        else if (sc->it >= sc->maxit) { sc->status = GRMP_ERR_NOCONV; sc->done = 1; }
    } else if (W == RED_VERIFY) {  // a = |b - A x|^2
        const double tr = sqrt(a / sc->rr0);
        sc->rr = a;
        sc->relres = tr;
        if (!isfinite(a)) { sc->status = GRMP_ERR_NONFINITE; sc->done = 1; }
        else if (tr <= sc->rtol) sc->done = 1;
        else if (sc->restarts >= 3 || (sc->restarts > 0 && tr > 0.5 * sc->best)) {
            // the true residual stopped improving: the FP64 floor of b - A x
            // (1e-13 .. 1e-10 relative for these problems) is accepted within
            // 100 rtol, as the engine does
            if (!(tr <= 100.0 * sc->rtol)) sc->status = GRMP_ERR_NOCONV;
            sc->done = 1;
        } else {
            sc->restarts += 1;
            sc->best = sc->restarts == 1 || tr < sc->best ? tr : sc->best;
        }
    } else if (W == RED_RESTART) {  // a = r.z of the true residual: p = z
        if (!(a > 0.0) || !isfinite(a)) { sc->status = GRMP_ERR_NONFINITE; sc->done = 1; return; }
        sc->rz = a;
        sc->beta = 0.0;
    } else {
        if (!isfinite(a)) { sc->status = GRMP_ERR_NONFINITE; sc->done = 1; return; }
        sc->beta = a / sc->rz;
        sc->rz = a;
    }
}


// ------------------------------------------------------------ row kernels
// The V-cycle's row products (RESID, CHEB_*) read FP32 copies of the level
// values and accumulate in FP64: the preconditioner moves 8 instead of 12
// bytes per nonzero.  The CG product (PCG_Q) and the true residual of the
// convergence check (TRUE_RESID) read the FP64 values, so the accepted
// solutions meet rtol against the FP64 operator.
enum MgOp { OP_RESID = 0, OP_CHEB_PRE = 1, OP_CHEB_POST1 = 2, OP_CHEB_POST2 = 3, OP_PCG_Q = 4, OP_TRUE_RESID = 5 };
template <int OP>
struct MgF64 {
    static constexpr bool value = OP == OP_PCG_Q || OP == OP_TRUE_RESID;
};

struct MgRows {
    const int32_t *rp, *ci;
    const double *val;
    const float *val32;  // FP32 copy (fixed rows 0) for the V-cycle's products
    const uint8_t *fixed;
    int64_t n;
    const double *v;     // the vector A multiplies
    const double *b, *dinv, *d;
    double *o1, *o2;
    const double *lam;   // Gershgorin bound of D^-1 A (device)
    const MgScal *sc;
    double *part;        // block partials (OP_PCG_Q: p.q; OP_CHEB_POST2 at level 0: r.z)
};

// 256 rows per block through the LDS-staged row product, then the epilogue:
//   RESID   o1 = b - A v
//   PRE     v = d0:  o1 = (1 + C1) d0 + C2 D^-1 (b - A d0)
//   POST1   v = x:   d = D^-1 (b - A x) / theta -> o1,  x + d -> o2
//   POST2   v = x1:  o1 = x1 + C1 d + C2 D^-1 (b - A x1)   [+ partials of b . o1]
//   PCG_Q   o1 = A v                                        [+ partials of v . o1]
// Fixed rows get 0 (no residual, no correction).
template <int OP>
__device__ __forceinline__ double mg_epilogue(int64_t i, double s, bool fx, const double *__restrict__ v,
                                              const double *__restrict__ b, const double *__restrict__ dinv,
                                              const double *__restrict__ d, const double *__restrict__ lam,
                                              double *__restrict__ o1, double *__restrict__ o2)
{
    if (OP == OP_RESID || OP == OP_TRUE_RESID) {
        o1[i] = fx ? 0.0 : b[i] - s;
    } else if (OP == OP_CHEB_PRE) {
        const double l = *lam;
        o1[i] = fx ? 0.0 : (1.0 + CHEB_C1) * v[i] + CHEB_C2(l) * dinv[i] * (b[i] - s);
    } else if (OP == OP_CHEB_POST1) {
        const double l = *lam;
        const double dd = fx ? 0.0 : dinv[i] * (b[i] - s) * CHEB_ITHETA(l);
        o1[i] = dd;
        o2[i] = v[i] + dd;
    } else if (OP == OP_CHEB_POST2) {
        const double l = *lam;
        const double z = fx ? 0.0 : v[i] + CHEB_C1 * d[i] + CHEB_C2(l) * dinv[i] * (b[i] - s);
        o1[i] = z;
        return b[i] * z;
    } else {
        const double q = fx ? 0.0 : s;
        o1[i] = q;
        return v[i] * q;
    }
    return 0.0;
}

template <int OP>
// Nonzeros staged per 256-row chunk: 2048 FP64 products (16 KB of LDS)
// lets twice as many workgroups share a CU as SPMV_CAP does; on the served
// stream that took the batched residual from 3.2 to 4.4 TB/s.  The sum order
// does not depend on it.
#define MG_ROWS_CAP (SPMV_CAP / 2)
__device__ __forceinline__ void mg_rows_body(const MgRows &a, int64_t r0, int pidx)
{
    __shared__ double lds[MG_ROWS_CAP];
    const int64_t r1 = r0 + KB < a.n ? r0 + KB : a.n;
    double s;
    if constexpr (MgF64<OP>::value) s = csr_rows_lds<double, double, double, MG_ROWS_CAP>(a.rp, a.ci, a.val, a.v, r0, r1, lds);
    else s = csr_rows_lds<double, float, double, MG_ROWS_CAP>(a.rp, a.ci, a.val32, a.v, r0, r1, lds);
    const int64_t i = r0 + threadIdx.x;
    double pa = 0.0, pb = 0.0;
    if (i < r1) pa = mg_epilogue<OP>(i, s, a.fixed[i] != 0, a.v, a.b, a.dinv, a.d, a.lam, a.o1, a.o2);
    if ((OP == OP_CHEB_POST2 || OP == OP_PCG_Q) && a.part) {
        block_sum2(pa, pb);
        if (threadIdx.x == 0) a.part[pidx] = pa;
    }
}

// Chebyshev pre-smoothing from x0 = 0 in one pass: d0 = D^-1 b / theta is
// formed on the fly in the staged row product (the same products as a stored
// d0), then o1 = (1 + C1) d0 + C2 D^-1 (b - A d0).
__device__ __forceinline__ void mg_pre_body(const MgRows &a, int64_t r0)
{
    __shared__ double lds[MG_ROWS_CAP];
    const double l = *a.lam, ith = CHEB_ITHETA(l);
    const int64_t r1 = r0 + KB < a.n ? r0 + KB : a.n;
    const int t = threadIdx.x;
    const int64_t r = r0 + t;
    const int32_t base = a.rp[r0], end = a.rp[r1];
    int32_t my0 = 0, my1 = 0;
    if (r < r1) {
        my0 = a.rp[r];
        my1 = a.rp[r + 1];
    }
    double s = 0.0;
    for (int32_t c0 = base; c0 < end; c0 += MG_ROWS_CAP) {
        const int32_t c1 = c0 + MG_ROWS_CAP < end ? c0 + MG_ROWS_CAP : end;
        __syncthreads();
        for (int32_t k = c0 + t; k < c1; k += KB) {
            const int32_t c = a.ci[k];
            lds[k - c0] = (double)a.val32[k] * (a.fixed[c] ? 0.0 : a.dinv[c] * a.b[c] * ith);
        }
        __syncthreads();
        const int32_t lo = my0 > c0 ? my0 : c0, hi = my1 < c1 ? my1 : c1;
        for (int32_t k = lo; k < hi; ++k) s += lds[k - c0];
    }
    if (r < r1) {
        const bool fx = a.fixed[r] != 0;
        const double d0 = fx ? 0.0 : a.dinv[r] * a.b[r] * ith;
        a.o1[r] = fx ? 0.0 : (1.0 + CHEB_C1) * d0 + CHEB_C2(l) * a.dinv[r] * (a.b[r] - s);
    }
}

__global__ __launch_bounds__(KB) void k_mg_pre(MgRows a)
{
    if (a.sc->done) return;
    mg_pre_body(a, (int64_t)blockIdx.x * KB);
}

template <int OP>
__global__ __launch_bounds__(KB) void k_mg_rows(MgRows a)
{
    if (a.sc->done) return;
    mg_rows_body<OP>(a, (int64_t)blockIdx.x * KB, blockIdx.x);
}

// the residual on levels of at least MG_FINE_ROWS rows: the same rows as
// k_mg_rows<OP_RESID> under its own name, the kernel the profiled runs time
#define MG_FINE_ROWS (1 << 21)
__global__ __launch_bounds__(KB) void k_mg_resid_fine(MgRows a)
{
    if (a.sc->done) return;
    mg_rows_body<OP_RESID>(a, (int64_t)blockIdx.x * KB, blockIdx.x);
}

// the V-cycle's FP32 copy of a level's values; fixed rows 0 (their 1e60
// penalty has no FP32 value, and their products are not used)
__global__ void k_mg_val32(const int32_t *__restrict__ rp, const double *__restrict__ val,
                           const uint8_t *__restrict__ fixed, float *__restrict__ val32, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool fx = fixed[i] != 0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) val32[k] = fx ? 0.f : (float)val[k];
}

// Gershgorin bound max_i sum_j |A_ij| / A_ii over free rows (a max, so the
// result does not depend on the order of the atomics)
__global__ void k_mg_gersh(const int32_t *__restrict__ rp, const double *__restrict__ val,
                           const double *__restrict__ dinv, const uint8_t *__restrict__ fixed, int64_t n,
                           unsigned long long *lam)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n && !fixed[i]) {
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += fabs(val[k]);
        s *= fabs(dinv[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_down(s, o, 64));
    if ((threadIdx.x & 63) == 0 && s > 0.0) atomicMax(lam, (unsigned long long)__double_as_longlong(s));
}

// ------------------------------------------------------------- transfers
// Fine node (i, j) of the Nf-square lies in coarse square (ic, jc) of the
// Nc-triangulation at offsets u = un / Nf, v = vn / Nf in [0, 1] (integers:
// i Nc = ic Nf + un), in its lower triangle when u >= v.  The interpolation
// weight of coarse node (ic + dx, jc + dy) there is w / Nf with
//   (0,0): Nf - max(un, vn)    (1,0): max(un - vn, 0)
//   (1,1): min(un, vn)         (0,1): max(vn - un, 0)
// (the barycentric coordinates of both triangles in one formula); prolongation
// and restriction both evaluate it, so P^T is exactly the transpose of P.
__device__ __forceinline__ void p1_coord(int Nf, int Nc, int i, int &ic, int &un)
{
    ic = i * Nc / Nf;
    if (ic > Nc - 1) ic = Nc - 1;
    un = i * Nc - ic * Nf;
}

__device__ __forceinline__ int p1_weight(int Nf, int dx, int dy, int un, int vn)
{
    if (dx == 0 && dy == 0) return Nf - (un > vn ? un : vn);
    if (dx == 1 && dy == 0) return un > vn ? un - vn : 0;
    if (dx == 1 && dy == 1) return un < vn ? un : vn;
    return vn > un ? vn - un : 0;  // (0,1)
}

// xf += P xc on free fine dofs (components blocked by node count); t < nnf * ncomp
__device__ __forceinline__ void prolong_p1_one(int64_t Nf64, int64_t Nc64, int ncomp, const double *__restrict__ xc,
                                               const uint8_t *__restrict__ fixed_f, double *__restrict__ xf, int64_t t)
{
    const int Nf = (int)Nf64, Nc = (int)Nc64;
    const int64_t nnf = (int64_t)(Nf + 1) * (Nf + 1), nnc = (int64_t)(Nc + 1) * (Nc + 1);
    if (fixed_f[t]) return;
    const int k = (int)(t / nnf);
    const int node = (int)(t - k * nnf);
    int ic, un, jc, vn;
    p1_coord(Nf, Nc, node % (Nf + 1), ic, un);
    p1_coord(Nf, Nc, node / (Nf + 1), jc, vn);
    const double inv = 1.0 / (double)Nf;
    const double *x = xc + k * nnc + (int64_t)jc * (Nc + 1) + ic;
    double s = (double)p1_weight(Nf, 0, 0, un, vn) * inv * x[0];
    s += (double)p1_weight(Nf, 1, 0, un, vn) * inv * x[1];
    s += (double)p1_weight(Nf, 1, 1, un, vn) * inv * x[Nc + 2];
    s += (double)p1_weight(Nf, 0, 1, un, vn) * inv * x[Nc + 1];
    xf[t] += s;
}

// bc = P^T rf on free coarse dofs, 0 on fixed ones; t < nnc * ncomp.  Coarse
// node (I, J) gathers the fine nodes of the index box around it whose coarse
// square touches it (at most 7 x 7 for Nc = ceil(Nf / 2)), row by row.
__device__ __forceinline__ void restrict_p1_one(int64_t Nf64, int64_t Nc64, int ncomp, const double *__restrict__ rf,
                                                const uint8_t *__restrict__ fixed_c, double *__restrict__ bc,
                                                int64_t t)
{
    const int Nf = (int)Nf64, Nc = (int)Nc64;
    const int64_t nnf = (int64_t)(Nf + 1) * (Nf + 1), nnc = (int64_t)(Nc + 1) * (Nc + 1);
    if (fixed_c[t]) {
        bc[t] = 0.0;
        return;
    }
    const int k = (int)(t / nnc);
    const int node = (int)(t - k * nnc);
    const int I = node % (Nc + 1), J = node / (Nc + 1);
    const int ilo = I >= 1 ? (I - 1) * Nf / Nc : 0, jlo = J >= 1 ? (J - 1) * Nf / Nc : 0;
    int ihi = ((I + 1) * Nf + Nc - 1) / Nc, jhi = ((J + 1) * Nf + Nc - 1) / Nc;
    if (ihi > Nf) ihi = Nf;
    if (jhi > Nf) jhi = Nf;
    int dxs[8], uns[8];
    const int ni = ihi - ilo + 1 < 8 ? ihi - ilo + 1 : 8;
    for (int a = 0; a < ni; ++a) {
        int ic;
        p1_coord(Nf, Nc, ilo + a, ic, uns[a]);
        dxs[a] = I - ic;
    }
    const double inv = 1.0 / (double)Nf;
    const double *r = rf + k * nnf;
    double s = 0.0;
    for (int j = jlo; j <= jhi; ++j) {
        int jc, vn;
        p1_coord(Nf, Nc, j, jc, vn);
        const int dy = J - jc;
        if (dy < 0 || dy > 1) continue;
        for (int a = 0; a < ni; ++a) {
            const int dx = dxs[a];
            if (dx < 0 || dx > 1) continue;
            const int w = p1_weight(Nf, dx, dy, uns[a], vn);
            if (w != 0) s += (double)w * inv * r[(int64_t)j * (Nf + 1) + ilo + a];
        }
    }
    bc[t] = s;
}

// P1 -> P2 on the same N-square, added into x2: node dofs copy (t < nn),
// face dofs take the mean of their end nodes (t < nc: each face is handled by
// the one cell whose CellFaceSign is +1 for it, its first appearance)
__device__ __forceinline__ void prolong_p2_one(int64_t N, const int32_t *__restrict__ cells,
                                               const int8_t *__restrict__ fsign, const double *__restrict__ x1,
                                               const uint8_t *__restrict__ fixed2, double *__restrict__ x2, int64_t t)
{
    const int64_t nn = (N + 1) * (N + 1), nc = 2 * N * N;
    if (t < nn && !fixed2[t]) x2[t] += x1[t];
    if (t >= nc) return;
    int64_t f[3];
    sq_cellfaces(N, t, f);
    for (int k = 0; k < 3; ++k) {
        if (fsign[k * nc + t] <= 0) continue;
        const int64_t dof = nn + f[k];
        if (fixed2[dof]) continue;
        const int32_t a = cells[k * nc + t], b = cells[((k + 1) % 3) * nc + t];
        x2[dof] += 0.5 * (x1[a] + x1[b]);
    }
}

// its transpose (v < nn): node v gathers r2[v] and half of each incident
// face's entry (left / right horizontal, lower / upper vertical, lower-left /
// upper-right diagonal)
__device__ __forceinline__ void restrict_p2_one(int64_t N, const double *__restrict__ r2,
                                                const uint8_t *__restrict__ fixed1, double *__restrict__ b1, int64_t v)
{
    const int64_t nn = (N + 1) * (N + 1);
    if (fixed1[v]) {
        b1[v] = 0.0;
        return;
    }
    const int64_t i = v % (N + 1), j = v / (N + 1);
    const double *rf = r2 + nn;
    double s = 0.0;
    if (i > 0) s += rf[sq_H(N, i - 1, j)];
    if (i < N) s += rf[sq_H(N, i, j)];
    if (j > 0) s += rf[sq_V(N, i, j - 1)];
    if (j < N) s += rf[sq_V(N, i, j)];
    if (i > 0 && j > 0) s += rf[sq_D(N, i - 1, j - 1)];
    if (i < N && j < N) s += rf[sq_D(N, i, j)];
    b1[v] = r2[v] + 0.5 * s;
}

__global__ void k_mg_prolong_p1(int64_t Nf, int64_t Nc, int ncomp, const double *__restrict__ xc,
                                const uint8_t *__restrict__ fixed_f, double *__restrict__ xf, const MgScal *sc)
{
    if (sc->done) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (Nf + 1) * (Nf + 1) * ncomp) prolong_p1_one(Nf, Nc, ncomp, xc, fixed_f, xf, t);
}

__global__ void k_mg_restrict_p1(int64_t Nf, int64_t Nc, int ncomp, const double *__restrict__ rf,
                                 const uint8_t *__restrict__ fixed_c, double *__restrict__ bc, const MgScal *sc)
{
    if (sc->done) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (Nc + 1) * (Nc + 1) * ncomp) restrict_p1_one(Nf, Nc, ncomp, rf, fixed_c, bc, t);
}

__global__ void k_mg_prolong_p2(int64_t N, const int32_t *__restrict__ cells, const int8_t *__restrict__ fsign,
                                const double *__restrict__ x1, const uint8_t *__restrict__ fixed2,
                                double *__restrict__ x2, const MgScal *sc)
{
    if (sc->done) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    prolong_p2_one(N, cells, fsign, x1, fixed2, x2, t);
}

__global__ void k_mg_restrict_p2(int64_t N, const double *__restrict__ r2, const uint8_t *__restrict__ fixed1,
                                 double *__restrict__ b1, const MgScal *sc)
{
    if (sc->done) return;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < (N + 1) * (N + 1)) restrict_p2_one(N, r2, fixed1, b1, v);
}

// coarse coefficient: geometric mean of the fine cells whose centroid lies in
// the coarse cell (centroids in units of 1 / (3 Nf); same integer test as the
// fine-to-coarse location above).  A coarse cell without a fine centroid
// (not met for Nf <= 2048) takes the fine cell holding its own centroid.
__device__ __forceinline__ int64_t mg_cell_of(int64_t Nf, int64_t Nc, int64_t X, int64_t Y)
{
    // point (X, Y) / (3 Nf) in the Nc-triangulation
    int64_t ic = X * Nc / (3 * Nf), jc = Y * Nc / (3 * Nf);
    if (ic > Nc - 1) ic = Nc - 1;
    if (jc > Nc - 1) jc = Nc - 1;
    const int64_t un = X * Nc - ic * 3 * Nf, vn = Y * Nc - jc * 3 * Nf;
    return 2 * (jc * Nc + ic) + (un >= vn ? 0 : 1);
}

__global__ void k_mg_coarse_kappa(int64_t Nf, int64_t Nc, const double *__restrict__ kf, double *__restrict__ kc)
{
    const int64_t cc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cc >= 2 * Nc * Nc) return;
    const int64_t sq = cc >> 1, jc = sq / Nc, ic = sq % Nc;
    // fine squares overlapping the coarse square
    const int64_t ilo = ic * Nf / Nc, jlo = jc * Nf / Nc;
    int64_t ihi = ((ic + 1) * Nf + Nc - 1) / Nc, jhi = ((jc + 1) * Nf + Nc - 1) / Nc;
    if (ihi > Nf) ihi = Nf;
    if (jhi > Nf) jhi = Nf;
    double s = 0.0;
    int cnt = 0;
    for (int64_t j = jlo; j < jhi; ++j)
        for (int64_t i = ilo; i < ihi; ++i)
            for (int up = 0; up < 2; ++up) {
                const int64_t X = 3 * i + (up ? 1 : 2), Y = 3 * j + (up ? 2 : 1);
                if (mg_cell_of(Nf, Nc, X, Y) != cc) continue;
                s += log(kf[2 * (j * Nf + i) + up]);
                ++cnt;
            }
    if (cnt == 0) {
        // centroid of the coarse cell in units of 1 / (3 Nc), located among fine cells
        const int64_t Xc = 3 * ic + ((cc & 1) ? 1 : 2), Yc = 3 * jc + ((cc & 1) ? 2 : 1);
        kc[cc] = kf[mg_cell_of(Nc, Nf, Xc, Yc)];
        return;
    }
    kc[cc] = exp(s / cnt);
}

// ------------------------------------------------------ coarsest level
// Dense inverse of the coarsest matrix (fixed rows and columns replaced by
// the identity), in-place Gauss-Jordan in LDS by one block.  The matrix is
// symmetric positive definite, so no pivoting is needed.
__global__ __launch_bounds__(256) void k_mg_dense_inv(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                      const double *__restrict__ val, const uint8_t *__restrict__ fixed,
                                                      int n, double *__restrict__ inv)
{
    __shared__ double M[MG_DENSE_MAX * MG_DENSE_MAX];
    __shared__ double f[MG_DENSE_MAX];
    const int t = threadIdx.x;
    for (int e = t; e < n * n; e += blockDim.x) M[e] = 0.0;
    __syncthreads();
    if (t < n) {
        if (fixed[t]) {
            M[t * n + t] = 1.0;
        } else {
            for (int32_t k = rp[t]; k < rp[t + 1]; ++k)
                if (!fixed[ci[k]]) M[t * n + ci[k]] = val[k];
        }
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double p = M[k * n + k];
        __syncthreads();
        if (t < n) M[k * n + t] = (t == k ? 1.0 : M[k * n + t]) / p;
        __syncthreads();
        if (t < n) f[t] = t == k ? 0.0 : M[t * n + k];
        __syncthreads();
        for (int e = t; e < n * n; e += blockDim.x) {
            const int r = e / n, c = e - r * n;
            if (r == k) continue;
            M[e] = (c == k ? 0.0 : M[e]) - f[r] * M[k * n + c];
        }
        __syncthreads();
    }
    for (int e = t; e < n * n; e += blockDim.x) inv[e] = M[e];
}

__global__ void k_mg_dense_apply(const double *__restrict__ inv, const double *__restrict__ b, int n,
                                 double *__restrict__ x, const MgScal *sc)
{
    if (sc->done) return;
    const int i = threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += inv[i * n + j] * b[j];
    x[i] = s;
}

// --------------------------------------------------------------- PCG
__global__ void k_mg_dot(const double *__restrict__ a, const double *__restrict__ b, int64_t n, double *part,
                         const MgScal *sc)
{
    if (sc->done) return;
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    double pa = i < n ? a[i] * b[i] : 0.0, pb = 0.0;
    block_sum2(pa, pb);
    if (threadIdx.x == 0) part[blockIdx.x] = pa;
}

// x += alpha p, r -= alpha q; partials of r.r
__global__ void k_mg_axpy(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                          const double *__restrict__ q, int64_t n, double *part, const MgScal *sc)
{
    if (sc->done) return;
    const double al = sc->alpha;
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    double pa = 0.0, pb = 0.0;
    if (i < n) {
        x[i] += al * p[i];
        const double ri = r[i] - al * q[i];
        r[i] = ri;
        pa = ri * ri;
    }
    block_sum2(pa, pb);
    if (threadIdx.x == 0) part[blockIdx.x] = pa;
}

__global__ void k_mg_update_p(double *__restrict__ p, const double *__restrict__ z, int64_t n, const MgScal *sc)
{
    if (sc->done) return;
    const double be = sc->beta;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + be * p[i];
}

// one block: the fixed-order sum of nb partials, then the scalar step
template <int W>
__global__ __launch_bounds__(KB) void k_mg_red(const double *__restrict__ part, int nb, const double *__restrict__ part2,
                                               MgScal *sc)
{
    if (sc->done) return;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += KB) {
        a += part[k];
        if (W == RED_INIT) b += part2[k];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) mg_scalar<W>(sc, a, b);
}

// a check was requested: resume the kernels for it
__global__ void k_mg_verify_begin(MgScal *sc)
{
    if (threadIdx.x == 0 && sc->done == 2) sc->done = 0;
}

// ------------------------------------------------- one-workgroup kernels
// Levels with at most block_rows rows run inside one workgroup of MG_BT
// threads, phases separated by workgroup barriers: the coarse end of every
// V-cycle in one launch, and for requests whose finest level is that small,
// whole CG iterations in one launch.  Rows are summed in column order by one
// thread each.
#define MG_BT 1024

struct MgBLevel {
    const int32_t *rp, *ci;
    const double *val, *dinv;
    const float *val32;
    const uint8_t *fixed;
    double *x, *b, *r, *d;
    const double *lam;
    int64_t n;
    int N, fe, ncomp, pad;
    const int32_t *cells;  // P2 level: its mesh, for the P1 -> P2 prolongation
    const int8_t *fsign;
};

struct MgBDesc {
    int first, nlev, nb0, pad;  // levels first .. nlev-1 run in the workgroup
    MgBLevel lv[MG_MAX_LEVELS];
    const double *inv;          // coarsest level's dense inverse
    double *x, *r, *z, *p, *q;  // level-0 CG vectors
    const double *rhs;
    double *part;               // level-0 block partials (nb0 entries)
    MgScal *sc;
};

// sum over the workgroup, returned to every thread
__device__ double blk_sum(double v)
{
    __shared__ double ws[MG_BT / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < MG_BT / 64; ++k) t += ws[k];
    return t;
}

// rows of level L through the epilogue OP; returns this thread's partial
template <int OP>
__device__ double blk_rows(const MgBLevel &L, const double *v, const double *b, const double *d, double *o1,
                           double *o2)
{
    double part = 0.0;
    for (int64_t i = threadIdx.x; i < L.n; i += MG_BT) {
        double s = 0.0;
        if constexpr (MgF64<OP>::value)
            for (int32_t k = L.rp[i]; k < L.rp[i + 1]; ++k) s += L.val[k] * v[L.ci[k]];
        else
            for (int32_t k = L.rp[i]; k < L.rp[i + 1]; ++k) s += (double)L.val32[k] * v[L.ci[k]];
        part += mg_epilogue<OP>(i, s, L.fixed[i] != 0, v, b, L.dinv, d, L.lam, o1, o2);
    }
    return part;
}

// V-cycle over levels D.first .. D.nlev-1 (x_first = V(b_first)); returns the
// workgroup total of b_0 . x_0 when D.first == 0
__device__ double blk_vcycle(const MgBDesc &D)
{
    const int L1 = D.nlev - 1;
    for (int l = D.first; l < L1; ++l) {
        const MgBLevel &L = D.lv[l], &C = D.lv[l + 1];
        for (int64_t i = threadIdx.x; i < L.n; i += MG_BT)
            L.d[i] = L.fixed[i] ? 0.0 : L.dinv[i] * L.b[i] * CHEB_ITHETA(*L.lam);
        __syncthreads();
        blk_rows<OP_CHEB_PRE>(L, L.d, L.b, nullptr, L.x, nullptr);
        __syncthreads();
        blk_rows<OP_RESID>(L, L.x, L.b, nullptr, L.r, nullptr);
        __syncthreads();
        for (int64_t t = threadIdx.x; t < C.n; t += MG_BT) {
            if (L.fe == GRMP_FE_H1P2) restrict_p2_one(L.N, L.r, C.fixed, C.b, t);
            else restrict_p1_one(L.N, C.N, L.ncomp, L.r, C.fixed, C.b, t);
        }
        __syncthreads();
    }
    {
        const MgBLevel &C = D.lv[L1];
        const int n = (int)C.n;
        for (int i = threadIdx.x; i < n; i += MG_BT) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += D.inv[i * n + j] * C.b[j];
            C.x[i] = s;
        }
        __syncthreads();
    }
    double rz = 0.0;
    for (int l = L1 - 1; l >= D.first; --l) {
        const MgBLevel &L = D.lv[l], &C = D.lv[l + 1];
        if (L.fe == GRMP_FE_H1P2) {
            const int64_t nn = (int64_t)(L.N + 1) * (L.N + 1), nc = 2ll * L.N * L.N;
            for (int64_t t = threadIdx.x; t < (nn > nc ? nn : nc); t += MG_BT)
                prolong_p2_one(L.N, L.cells, L.fsign, C.x, L.fixed, L.x, t);
        } else {
            for (int64_t t = threadIdx.x; t < L.n; t += MG_BT)
                prolong_p1_one(L.N, C.N, L.ncomp, C.x, L.fixed, L.x, t);
        }
        __syncthreads();
        blk_rows<OP_CHEB_POST1>(L, L.x, L.b, nullptr, L.d, L.r);
        __syncthreads();
        const double p = blk_rows<OP_CHEB_POST2>(L, L.r, L.b, L.d, L.x, nullptr);
        __syncthreads();
        if (l == 0) rz = blk_sum(p);
    }
    if (D.first == 0 && L1 == 0) {  // a single level: z is the exact solve
        double p = 0.0;
        for (int64_t i = threadIdx.x; i < D.lv[0].n; i += MG_BT) p += D.lv[0].b[i] * D.lv[0].x[i];
        rz = blk_sum(p);
    }
    return rz;
}

// coarse end of a V-cycle (levels D.first ..); at D.first == 0 also the
// level-0 partials of r.z (total in part[0], zeros after it)
__global__ __launch_bounds__(MG_BT) void k_mg_vcycle_block(const MgBDesc *__restrict__ Dp, const MgScal *sc)
{
    if (sc->done) return;
    const MgBDesc &D = *Dp;
    const double rz = blk_vcycle(D);
    if (D.first == 0)
        for (int k = threadIdx.x; k < D.nb0; k += MG_BT) D.part[k] = k == 0 ? rz : 0.0;
}

// up to iters CG iterations of a request whose levels all fit the workgroup,
// with the true-residual check and restart of mg_job_verify done in place
__device__ void mg_pcg_block_body(const MgBDesc &D, MgScal *sc, int iters)
{
    const MgBLevel &L0 = D.lv[0];
    const int64_t n = L0.n;
    for (int k = 0; k < iters; ++k) {
        __syncthreads();
        const int st = sc->done;
        __syncthreads();
        if (st == 1) return;
        if (st == 2) {  // the updated residual met rtol: check b - A x
            blk_rows<OP_TRUE_RESID>(L0, D.x, D.rhs, nullptr, D.r, nullptr);
            __syncthreads();
            double p = 0.0;
            for (int64_t i = threadIdx.x; i < n; i += MG_BT) p += D.r[i] * D.r[i];
            const double rr = blk_sum(p);
            if (threadIdx.x == 0) {
                sc->done = 0;
                mg_scalar<RED_VERIFY>(sc, rr, 0.0);
            }
            __syncthreads();
            if (sc->done) continue;
            const double rz = blk_vcycle(D);
            if (threadIdx.x == 0) mg_scalar<RED_RESTART>(sc, rz, 0.0);
            __syncthreads();
            if (sc->done) continue;
            for (int64_t i = threadIdx.x; i < n; i += MG_BT) D.p[i] = D.z[i];
            continue;
        }
        const double pq = blk_sum(blk_rows<OP_PCG_Q>(L0, D.p, nullptr, nullptr, D.q, nullptr));
        if (threadIdx.x == 0) mg_scalar<RED_ALPHA>(sc, pq, 0.0);
        __syncthreads();
        if (sc->done) continue;
        const double al = sc->alpha;
        double p = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += MG_BT) {
            D.x[i] += al * D.p[i];
            const double ri = D.r[i] - al * D.q[i];
            D.r[i] = ri;
            p += ri * ri;
        }
        const double rr = blk_sum(p);
        if (threadIdx.x == 0) mg_scalar<RED_CONV>(sc, rr, 0.0);
        __syncthreads();
        if (sc->done) continue;
        const double rz = blk_vcycle(D);
        if (threadIdx.x == 0) mg_scalar<RED_BETA>(sc, rz, 0.0);
        __syncthreads();
        if (sc->done) continue;
        const double be = sc->beta;
        for (int64_t i = threadIdx.x; i < n; i += MG_BT) D.p[i] = D.z[i] + be * D.p[i];
    }
}

__global__ __launch_bounds__(MG_BT) void k_mg_pcg_block(const MgBDesc *__restrict__ Dp, MgScal *sc, int iters)
{
    mg_pcg_block_body(*Dp, sc, iters);
}

// ------------------------------------------------------ lockstep batches
// Requests in flight advance together: every phase of a CG iteration and of
// its V-cycle is one launch over a table of 256-row tiles of all requests
// taking part, each tile naming its request and level.  Requests whose
// levels all fit one workgroup run in one launch, a workgroup each.  The
// reductions stay per request (one block each, the same fixed order), so a
// request's iterates are the ones it has alone.
struct MgTile {
    int32_t req, level, row0, pad;
};

// a tile is MG_TILE_CHUNKS consecutive 256-row chunks of one level, handled
// by one workgroup chunk after chunk: the tile and descriptor lookups that
// start a workgroup are paid once per tile.  Partials stay per 256-row chunk.
// Two measured best for the batched residual on the served stream (217 us
// per launch against 223-228 us with four and 222-225 us with one).
#ifndef MG_TILE_CHUNKS
#define MG_TILE_CHUNKS 2
#endif

__device__ __forceinline__ MgRows mg_tile_rows(const MgBDesc &D, int level)
{
    const MgBLevel &L = D.lv[level];
    MgRows a;
    a.rp = L.rp;
    a.ci = L.ci;
    a.val = L.val;
    a.val32 = L.val32;
    a.fixed = L.fixed;
    a.n = L.n;
    a.v = nullptr;
    a.b = nullptr;
    a.dinv = L.dinv;
    a.d = nullptr;
    a.o1 = nullptr;
    a.o2 = nullptr;
    a.lam = L.lam;
    a.sc = D.sc;
    a.part = nullptr;
    return a;
}

template <int OP>
__global__ __launch_bounds__(KB) void k_mgb_rows(const MgTile *__restrict__ T, const MgBDesc *const *__restrict__ Dp)
{
    const MgTile t = T[blockIdx.x];
    const MgBDesc &D = *Dp[t.req];
    if (D.sc->done) return;
    const MgBLevel &L = D.lv[t.level];
    MgRows a = mg_tile_rows(D, t.level);
    if (OP == OP_RESID) {
        a.v = L.x;
        a.b = L.b;
        a.o1 = L.r;
    } else if (OP == OP_CHEB_PRE) {
        a.v = L.d;
        a.b = L.b;
        a.o1 = L.x;
    } else if (OP == OP_CHEB_POST1) {
        a.v = L.x;
        a.b = L.b;
        a.o1 = L.d;
        a.o2 = L.r;
    } else if (OP == OP_CHEB_POST2) {
        a.v = L.r;
        a.b = L.b;
        a.d = L.d;
        a.o1 = L.x;
        a.part = t.level == 0 ? D.part : nullptr;
    } else if (OP == OP_PCG_Q) {
        a.v = D.p;
        a.o1 = D.q;
        a.part = D.part;
    }
    for (int c = 0; c < MG_TILE_CHUNKS; ++c) {
        const int64_t r0 = t.row0 + (int64_t)c * KB;
        if (r0 >= L.n) break;
        mg_rows_body<OP>(a, r0, (int)(r0 / KB));
    }
}

// d0 = D^-1 b / theta on the level's rows, the first step of the
// pre-smoothing, stored for k_mgb_rows<OP_CHEB_PRE>.  Batched launches cost
// little, so the separate pass (about a sixth of the product's bytes) beats
// forming d0 inside the product, which gathers three arrays per nonzero.
__global__ __launch_bounds__(KB) void k_mgb_first(const MgTile *__restrict__ T, const MgBDesc *const *__restrict__ Dp)
{
    const MgTile t = T[blockIdx.x];
    const MgBDesc &D = *Dp[t.req];
    if (D.sc->done) return;
    const MgBLevel &L = D.lv[t.level];
    const double ith = CHEB_ITHETA(*L.lam);
    for (int c = 0; c < MG_TILE_CHUNKS; ++c) {
        const int64_t i = t.row0 + (int64_t)c * KB + threadIdx.x;
        if (i < L.n) L.d[i] = L.fixed[i] ? 0.0 : L.dinv[i] * L.b[i] * ith;
    }
}

// tiles over the rows of coarse level t.level (> 0)
__global__ __launch_bounds__(KB) void k_mgb_restrict(const MgTile *__restrict__ T,
                                                     const MgBDesc *const *__restrict__ Dp)
{
    const MgTile t = T[blockIdx.x];
    const MgBDesc &D = *Dp[t.req];
    if (D.sc->done) return;
    const MgBLevel &F = D.lv[t.level - 1], &C = D.lv[t.level];
    for (int c = 0; c < MG_TILE_CHUNKS; ++c) {
        const int64_t i = t.row0 + (int64_t)c * KB + threadIdx.x;
        if (i >= C.n) break;
        if (F.fe == GRMP_FE_H1P2) restrict_p2_one(F.N, F.r, C.fixed, C.b, i);
        else restrict_p1_one(F.N, C.N, F.ncomp, F.r, C.fixed, C.b, i);
    }
}

// items of fine level t.level: its rows, or for P2 max(nodes, cells)
__global__ __launch_bounds__(KB) void k_mgb_prolong(const MgTile *__restrict__ T,
                                                    const MgBDesc *const *__restrict__ Dp)
{
    const MgTile t = T[blockIdx.x];
    const MgBDesc &D = *Dp[t.req];
    if (D.sc->done) return;
    const MgBLevel &F = D.lv[t.level], &C = D.lv[t.level + 1];
    for (int c = 0; c < MG_TILE_CHUNKS; ++c) {
        const int64_t i = t.row0 + (int64_t)c * KB + threadIdx.x;
        if (F.fe == GRMP_FE_H1P2) {
            const int64_t nn = (int64_t)(F.N + 1) * (F.N + 1), nc = 2ll * F.N * F.N;
            if (i < (nn > nc ? nn : nc)) prolong_p2_one(F.N, F.cells, F.fsign, C.x, F.fixed, F.x, i);
        } else if (i < F.n) {
            prolong_p1_one(F.N, C.N, F.ncomp, C.x, F.fixed, F.x, i);
        }
    }
}

// level-0 tiles: x += alpha p, r -= alpha q, partials of r.r
__global__ __launch_bounds__(KB) void k_mgb_axpy(const MgTile *__restrict__ T, const MgBDesc *const *__restrict__ Dp)
{
    const MgTile t = T[blockIdx.x];
    const MgBDesc &D = *Dp[t.req];
    if (D.sc->done) return;
    const double al = D.sc->alpha;
    const int64_t n = D.lv[0].n;
    for (int c = 0; c < MG_TILE_CHUNKS; ++c) {
        const int64_t r0 = t.row0 + (int64_t)c * KB;
        if (r0 >= n) break;
        const int64_t i = r0 + threadIdx.x;
        double pa = 0.0, pb = 0.0;
        if (i < n) {
            D.x[i] += al * D.p[i];
            const double ri = D.r[i] - al * D.q[i];
            D.r[i] = ri;
            pa = ri * ri;
        }
        block_sum2(pa, pb);
        if (threadIdx.x == 0) D.part[r0 / KB] = pa;
    }
}

__global__ __launch_bounds__(KB) void k_mgb_update_p(const MgTile *__restrict__ T,
                                                     const MgBDesc *const *__restrict__ Dp)
{
    const MgTile t = T[blockIdx.x];
    const MgBDesc &D = *Dp[t.req];
    if (D.sc->done) return;
    const double be = D.sc->beta;
    for (int c = 0; c < MG_TILE_CHUNKS; ++c) {
        const int64_t i = t.row0 + (int64_t)c * KB + threadIdx.x;
        if (i < D.lv[0].n) D.p[i] = D.z[i] + be * D.p[i];
    }
}

// one block per request: the fixed-order sum of its partials, then step W
template <int W>
__global__ __launch_bounds__(KB) void k_mgb_red(const MgBDesc *const *__restrict__ Dp, const int32_t *__restrict__ reqs)
{
    const MgBDesc &D = *Dp[reqs[blockIdx.x]];
    if (D.sc->done) return;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < D.nb0; k += KB) a += D.part[k];
    block_sum2(a, b);
    if (threadIdx.x == 0) mg_scalar<W>(D.sc, a, b);
}

__global__ __launch_bounds__(MG_BT) void k_mgb_vcycle_block(const MgBDesc *const *__restrict__ Dp,
                                                            const int32_t *__restrict__ reqs)
{
    const MgBDesc &D = *Dp[reqs[blockIdx.x]];
    if (D.sc->done) return;
    blk_vcycle(D);
}

__global__ __launch_bounds__(MG_BT) void k_mgb_pcg_block(const MgBDesc *const *__restrict__ Dp,
                                                         const int32_t *__restrict__ reqs, int iters)
{
    const MgBDesc &D = *Dp[reqs[blockIdx.x]];
    mg_pcg_block_body(D, D.sc, iters);
}

__global__ void k_mgb_status(const MgBDesc *const *__restrict__ Dp, int n, MgStatus *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MgScal *sc = Dp[i]->sc;
    out[i].relres = sc->relres;
    out[i].it = sc->it;
    out[i].done = sc->done;
    out[i].status = sc->status;
    out[i].pad = 0;
}

// ------------------------------------------------------------- the job
struct MgLevel {
    MgLevelIn in;
    int64_t n = 0, nnz = 0, nc = 0;
    double *kappa = nullptr, *val = nullptr, *dinv = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;
    float *dinv32 = nullptr;
    float *val32 = nullptr;  // FP32 copy of val for the V-cycle, fixed rows 0
    double *lam = nullptr;
    DevForm *dforms = nullptr;
    DevForm hforms[4];
};

struct MgJob {
    grmp_ctx *ctx = nullptr;
    hipStream_t st = nullptr;
    int nlev = 0;
    MgLevel lv[MG_MAX_LEVELS];
    char *buf = nullptr;
    size_t bytes = 0;
    double *x = nullptr, *rhs = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double *part = nullptr, *part2 = nullptr, *inv = nullptr;
    MgScal *sc = nullptr;
    int nb0 = 0;
    int first_small = 0;      // first level that runs inside one workgroup
    MgBDesc hdesc;
    MgBDesc *ddesc = nullptr;
    hipEvent_t ev = nullptr;
    // grid-path chunk of CG iterations captured once and replayed
    bool use_graph = false;
    int graph_iters = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // profile: event pairs around every OP_RESID launch, with its algorithmic bytes
    bool profile = false;
    std::vector<hipEvent_t> pev, pool;  // profiled launches in flight; free events
    std::vector<double> pbytes;
    double prof_s = 0.0, prof_bytes = 0.0;
    int64_t prof_launches = 0;
};

static inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t mg_layout(MgJob *j, bool assign)
{
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += al256(bytes); return o; };
    const int64_t n0 = j->lv[0].n;
    const int nb0 = (int)nblk(n0, KB);
    size_t o_x = take(n0 * 8), o_rhs = take(n0 * 8), o_r = take(n0 * 8), o_z = take(n0 * 8), o_p = take(n0 * 8),
           o_q = take(n0 * 8), o_part = take(nb0 * 8), o_part2 = take(nb0 * 8), o_sc = take(sizeof(MgScal)),
           o_desc = take(sizeof(MgBDesc));
    const MgLevel &c = j->lv[j->nlev - 1];
    size_t o_inv = take(c.n * c.n * 8);
    std::vector<size_t> o(MG_MAX_LEVELS * 10);
    for (int l = 0; l < j->nlev; ++l) {
        const MgLevel &L = j->lv[l];
        const bool own_kappa = l > 0 && L.in.N != j->lv[l - 1].in.N;
        o[10 * l + 0] = own_kappa || l == 0 ? take(L.nc * 8) : 0;
        o[10 * l + 1] = take(L.nnz * 8);
        o[10 * l + 2] = take(L.n * 8);
        o[10 * l + 3] = take(L.n * 4);
        o[10 * l + 4] = l == 0 ? 0 : take(L.n * 8);  // x (level 0: z)
        o[10 * l + 5] = l == 0 ? 0 : take(L.n * 8);  // b (level 0: r)
        o[10 * l + 6] = take(L.n * 8);
        o[10 * l + 7] = take(L.n * 8);
        o[10 * l + 8] = take(8 + sizeof(DevForm) * 4);
        o[10 * l + 9] = take(L.nnz * 4);
    }
    if (!assign) return off;
    char *B = j->buf;
    j->x = (double *)(B + o_x);
    j->rhs = (double *)(B + o_rhs);
    j->r = (double *)(B + o_r);
    j->z = (double *)(B + o_z);
    j->p = (double *)(B + o_p);
    j->q = (double *)(B + o_q);
    j->part = (double *)(B + o_part);
    j->part2 = (double *)(B + o_part2);
    j->sc = (MgScal *)(B + o_sc);
    j->ddesc = (MgBDesc *)(B + o_desc);
    j->inv = (double *)(B + o_inv);
    j->nb0 = nb0;
    for (int l = 0; l < j->nlev; ++l) {
        MgLevel &L = j->lv[l];
        const bool own_kappa = l > 0 && L.in.N != j->lv[l - 1].in.N;
        L.kappa = (own_kappa || l == 0) ? (double *)(B + o[10 * l + 0]) : j->lv[l - 1].kappa;
        L.val = (double *)(B + o[10 * l + 1]);
        L.dinv = (double *)(B + o[10 * l + 2]);
        L.dinv32 = (float *)(B + o[10 * l + 3]);
        L.x = l == 0 ? j->z : (double *)(B + o[10 * l + 4]);
        L.b = l == 0 ? j->r : (double *)(B + o[10 * l + 5]);
        L.r = (double *)(B + o[10 * l + 6]);
        L.d = (double *)(B + o[10 * l + 7]);
        L.lam = (double *)(B + o[10 * l + 8]);
        L.dforms = (DevForm *)(B + o[10 * l + 8] + 8);
        L.val32 = (float *)(B + o[10 * l + 9]);
    }
    return off;
}

int64_t mg_job_bytes(const MgLevelIn *in, int nlev)
{
    MgJob j;
    j.nlev = nlev;
    for (int l = 0; l < nlev; ++l) {
        j.lv[l].in = in[l];
        j.lv[l].n = in[l].A->n;
        j.lv[l].nnz = in[l].A->nnz;
        j.lv[l].nc = in[l].sp->mesh->ncells;
    }
    return (int64_t)mg_layout(&j, false);
}

static void fill_desc(MgJob *j)
{
    MgBDesc &D = j->hdesc;
    memset(&D, 0, sizeof(D));
    D.first = j->first_small;
    D.nlev = j->nlev;
    D.nb0 = j->nb0;
    for (int l = 0; l < j->nlev; ++l) {
        const MgLevel &L = j->lv[l];
        MgBLevel &B = D.lv[l];
        B.rp = L.in.A->rowptr;
        B.ci = L.in.A->cols;
        B.val = L.val;
        B.val32 = L.val32;
        B.dinv = L.dinv;
        B.fixed = L.in.A->fixed;
        B.x = L.x;
        B.b = L.b;
        B.r = L.r;
        B.d = L.d;
        B.lam = L.lam;
        B.n = L.n;
        B.N = L.in.N;
        B.fe = L.in.fe;
        B.ncomp = L.in.ncomp;
        B.cells = L.in.sp->mesh->cells;
        B.fsign = L.in.sp->mesh->fsign;
    }
    D.inv = j->inv;
    D.x = j->x;
    D.r = j->r;
    D.z = j->z;
    D.p = j->p;
    D.q = j->q;
    D.rhs = j->rhs;
    D.part = j->part;
    D.sc = j->sc;
}

int mg_job_create(grmp_ctx *ctx, hipStream_t st, const MgLevelIn *in, int nlev, bool profile, int block_rows,
                  int coarse_rows, bool graph, MgJob **out)
{
    GRMP_CHECK(ctx, nlev >= 1 && nlev <= MG_MAX_LEVELS, GRMP_ERR_ARG, "multigrid: %d levels", nlev);
    for (int l = 0; l < nlev; ++l) {
        const MgLevelIn &L = in[l];
        GRMP_CHECK(ctx, L.A && L.sp && L.A->have_pattern && L.sp->mesh->kind == MESH_UNITSQUARE && L.sp->mesh->N == L.N &&
                            L.A->n == L.sp->ndofs && L.A->n < (1ll << 31),
                   GRMP_ERR_ARG, "multigrid: level %d is not an assembled unit-square pattern", l);
        if (l > 0) {
            const MgLevelIn &F = in[l - 1];
            const bool p2 = F.fe == GRMP_FE_H1P2 && L.fe == GRMP_FE_H1P1 && L.N == F.N && L.ncomp == 1 && F.ncomp == 1;
            const bool p1 = F.fe == GRMP_FE_H1P1 && L.fe == GRMP_FE_H1P1 && L.N == (F.N + 1) / 2 && L.ncomp == F.ncomp;
            GRMP_CHECK(ctx, p2 || p1, GRMP_ERR_ARG, "multigrid: level %d does not coarsen level %d", l, l - 1);
        }
    }
    GRMP_CHECK(ctx, in[nlev - 1].A->n <= MG_DENSE_MAX, GRMP_ERR_ARG, "multigrid: coarsest level has %lld > %d dofs",
               (long long)in[nlev - 1].A->n, MG_DENSE_MAX);
    MgJob *j = new MgJob();
    j->ctx = ctx;
    j->st = st;
    j->nlev = nlev;
    j->profile = profile;
    j->use_graph = graph && !profile;  // events cannot be recorded inside a stream capture
    for (int l = 0; l < nlev; ++l) {
        j->lv[l].in = in[l];
        j->lv[l].n = in[l].A->n;
        j->lv[l].nnz = in[l].A->nnz;
        j->lv[l].nc = in[l].sp->mesh->ncells;
    }
    j->bytes = mg_layout(j, false);
    hipError_t e = hipMallocAsync((void **)&j->buf, j->bytes, st);
    if (e != hipSuccess) {
        delete j;
        return grmp_fail(ctx, e == hipErrorOutOfMemory ? GRMP_ERR_OOM : GRMP_ERR_HIP, "multigrid: device buffer");
    }
    mg_layout(j, true);
    // a request with at most block_rows rows runs whole in one workgroup;
    // otherwise its levels with at most coarse_rows rows do
    coarse_rows = std::max(coarse_rows, MG_DENSE_MAX);
    j->first_small = nlev - 1;
    if (j->lv[0].n <= block_rows) j->first_small = 0;
    else
        while (j->first_small > 1 && j->lv[j->first_small - 1].n <= coarse_rows) --j->first_small;
    fill_desc(j);
    if (hipEventCreateWithFlags(&j->ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipFreeAsync(j->buf, st);
        delete j;
        return grmp_fail(ctx, GRMP_ERR_HIP, "multigrid: event");
    }
    *out = j;
    return GRMP_OK;
}

double *mg_job_kappa0(MgJob *j) { return j->lv[0].kappa; }
double *mg_job_rhs(MgJob *j) { return j->rhs; }
const double *mg_job_solution(const MgJob *j) { return j->x; }

static void rows_args(MgRows &a, const MgLevel &L, const MgJob *j)
{
    memset(&a, 0, sizeof(a));
    a.rp = L.in.A->rowptr;
    a.ci = L.in.A->cols;
    a.val = L.val;
    a.val32 = L.val32;
    a.fixed = L.in.A->fixed;
    a.n = L.n;
    a.dinv = L.dinv;
    a.lam = L.lam;
    a.sc = j->sc;
}

// Residuals on levels of at least MG_FINE_ROWS rows run as k_mg_resid_fine;
// when profiled, each such launch carries a start / stop event pair bound to
// its own dispatch (hipExtLaunchKernelGGL), kept with its algorithmic bytes:
// values and columns, row pointers, x gathered once, b read, r written, mask.
template <int OP>
static void launch_rows(MgJob *j, const MgRows &a, int64_t nnz = 0)
{
    if (OP == OP_RESID && a.n >= MG_FINE_ROWS) {
        if (!j->profile) {
            hipLaunchKernelGGL(k_mg_resid_fine, dim3(nblk(a.n, KB)), dim3(KB), 0, j->st, a);
            return;
        }
        hipEvent_t e0, e1;
        if (!ev_take(j->pool, e0, e1)) {
            hipLaunchKernelGGL(k_mg_resid_fine, dim3(nblk(a.n, KB)), dim3(KB), 0, j->st, a);
            return;
        }
        hipExtLaunchKernelGGL(k_mg_resid_fine, dim3(nblk(a.n, KB)), dim3(KB), 0, j->st, e0, e1, 0, a);
        j->pev.push_back(e0);
        j->pev.push_back(e1);
        // FP32 values + int32 columns, row pointers, x gathered once, b, r, mask
        j->pbytes.push_back((double)nnz * 8.0 + 4.0 * (a.n + 1) + 24.0 * a.n + 1.0 * a.n);
        return;
    }
    hipLaunchKernelGGL(k_mg_rows<OP>, dim3(nblk(a.n, KB)), dim3(KB), 0, j->st, a);
}

// restriction of level l's residual r_l into b_{l+1}
static void mg_restrict(MgJob *j, int l)
{
    const MgLevel &F = j->lv[l], &C = j->lv[l + 1];
    if (F.in.fe == GRMP_FE_H1P2)
        hipLaunchKernelGGL(k_mg_restrict_p2, dim3(nblk(C.n, 256)), dim3(256), 0, j->st, (int64_t)F.in.N, F.r,
                           C.in.A->fixed, C.b, j->sc);
    else
        hipLaunchKernelGGL(k_mg_restrict_p1, dim3(nblk(C.n, 256)), dim3(256), 0, j->st, (int64_t)F.in.N,
                           (int64_t)C.in.N, F.in.ncomp, F.r, C.in.A->fixed, C.b, j->sc);
}

static void mg_prolong(MgJob *j, int l)
{
    const MgLevel &F = j->lv[l], &C = j->lv[l + 1];
    if (F.in.fe == GRMP_FE_H1P2) {
        const grmp_mesh *m = F.in.sp->mesh;
        const int64_t nn = m->nnodes, nc = m->ncells;
        hipLaunchKernelGGL(k_mg_prolong_p2, dim3(nblk(std::max(nn, nc), 256)), dim3(256), 0, j->st, (int64_t)F.in.N,
                           m->cells, m->fsign, C.x, F.in.A->fixed, F.x, j->sc);
    } else {
        hipLaunchKernelGGL(k_mg_prolong_p1, dim3(nblk(F.n, 256)), dim3(256), 0, j->st, (int64_t)F.in.N,
                           (int64_t)C.in.N, F.in.ncomp, C.x, F.in.A->fixed, F.x, j->sc);
    }
}

// x_0 (= z) = V(b_0 (= r)); the level-0 block partials of r.z land in part.
// Levels below first_small run as grid kernels, the rest in one workgroup.
static void mg_vcycle(MgJob *j)
{
    const int F = j->first_small;
    for (int l = 0; l < F; ++l) {
        MgLevel &L = j->lv[l];
        MgRows a;
        rows_args(a, L, j);
        a.b = L.b;
        a.o1 = L.x;
        hipLaunchKernelGGL(k_mg_pre, dim3(nblk(L.n, KB)), dim3(KB), 0, j->st, a);
        rows_args(a, L, j);
        a.v = L.x;
        a.b = L.b;
        a.o1 = L.r;
        launch_rows<OP_RESID>(j, a, L.nnz);
        mg_restrict(j, l);
    }
    hipLaunchKernelGGL(k_mg_vcycle_block, dim3(1), dim3(MG_BT), 0, j->st, j->ddesc, j->sc);
    for (int l = F - 1; l >= 0; --l) {
        MgLevel &L = j->lv[l];
        mg_prolong(j, l);
        MgRows a;
        rows_args(a, L, j);
        a.v = L.x;
        a.b = L.b;
        a.o1 = L.d;
        a.o2 = L.r;
        launch_rows<OP_CHEB_POST1>(j, a);
        rows_args(a, L, j);
        a.v = L.r;
        a.b = L.b;
        a.d = L.d;
        a.o1 = L.x;
        a.part = l == 0 ? j->part : nullptr;
        launch_rows<OP_CHEB_POST2>(j, a);
    }
}

// level values, smoother bounds and the coarsest inverse: once per operator
static int mg_job_assemble(MgJob *j)
{
    grmp_ctx *ctx = j->ctx;
    hipStream_t st = j->st;
    for (int l = 0; l < j->nlev; ++l) {
        MgLevel &L = j->lv[l];
        if (l > 0 && L.in.N != j->lv[l - 1].in.N)
            hipLaunchKernelGGL(k_mg_coarse_kappa, dim3(nblk(L.nc, 256)), dim3(256), 0, st, (int64_t)j->lv[l - 1].in.N,
                               (int64_t)L.in.N, j->lv[l - 1].kappa, L.kappa);
        GRMP_TRY(assemble_values(L.in.A, L.kappa, L.dforms, L.hforms, L.val, nullptr, L.dinv, L.dinv32, st));
        hipLaunchKernelGGL(k_mg_val32, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.in.A->rowptr, L.val, L.in.A->fixed,
                           L.val32, L.n);
        GRMP_HIP(ctx, hipMemsetAsync(L.lam, 0, 8, st));
        hipLaunchKernelGGL(k_mg_gersh, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.in.A->rowptr, L.val, L.dinv,
                           L.in.A->fixed, L.n, (unsigned long long *)L.lam);
    }
    const MgLevel &C = j->lv[j->nlev - 1];
    hipLaunchKernelGGL(k_mg_dense_inv, dim3(1), dim3(256), 0, st, C.in.A->rowptr, C.in.A->cols, C.val, C.in.A->fixed,
                       (int)C.n, j->inv);
    GRMP_HIP(ctx, hipMemcpyAsync(j->ddesc, &j->hdesc, sizeof(MgBDesc), hipMemcpyHostToDevice, st));
    return GRMP_OK;
}

// Krylov state for the right-hand side in mg_job_rhs: scalars reset,
// x = 0, r = rhs (0 on fixed rows), z = V(r), p = 0 then p = z
static int mg_job_start(MgJob *j, double rtol, int maxit)
{
    grmp_ctx *ctx = j->ctx;
    hipStream_t st = j->st;
    MgScal s0;
    memset(&s0, 0, sizeof(s0));
    s0.rtol = rtol;
    s0.maxit = maxit;
    GRMP_HIP(ctx, hipMemcpyAsync(j->sc, &s0, sizeof(s0), hipMemcpyHostToDevice, st));
    const int64_t n = j->lv[0].n;
    GRMP_HIP(ctx, hipMemsetAsync(j->x, 0, n * 8, st));
    GRMP_HIP(ctx, hipMemsetAsync(j->p, 0, n * 8, st));
    GRMP_HIP(ctx, hipMemcpyAsync(j->r, j->rhs, n * 8, hipMemcpyDeviceToDevice, st));
    GRMP_TRY(zero_fixed_launch(j->lv[0].in.A, j->r, st));
    mg_vcycle(j);
    hipLaunchKernelGGL(k_mg_dot, dim3(j->nb0), dim3(KB), 0, st, j->r, j->r, n, j->part2, j->sc);
    hipLaunchKernelGGL(k_mg_red<RED_INIT>, dim3(1), dim3(KB), 0, st, j->part, j->nb0, j->part2, j->sc);
    hipLaunchKernelGGL(k_mg_update_p, dim3(nblk(n, 256)), dim3(256), 0, st, j->p, j->z, n, j->sc);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

int mg_job_setup(MgJob *j, double rtol, int maxit)
{
    GRMP_TRY(mg_job_assemble(j));
    return mg_job_start(j, rtol, maxit);
}

// The setup of many jobs with their level assembly batched: coarse
// coefficients of every job, one launch assembling every level of every job
// (assemble_batch, the same values as mg_job_assemble's per-level launches),
// then per job the Gershgorin bounds, the coarsest inverse and the first
// V-cycle.  The jobs share the stream st.
int mg_jobs_setup_batched(MgJob *const *jobs, int n, hipStream_t st, double *const rtol, int maxit)
{
    if (n <= 0) return GRMP_OK;
    grmp_ctx *ctx = jobs[0]->ctx;
    std::vector<AsmBatchItem> items;
    for (int k = 0; k < n; ++k) {
        MgJob *j = jobs[k];
        GRMP_CHECK(ctx, j->st == st, GRMP_ERR_ARG, "batched multigrid setup: jobs on different streams");
        for (int l = 0; l < j->nlev; ++l) {
            MgLevel &L = j->lv[l];
            if (l > 0 && L.in.N != j->lv[l - 1].in.N)
                hipLaunchKernelGGL(k_mg_coarse_kappa, dim3(nblk(L.nc, 256)), dim3(256), 0, st, (int64_t)j->lv[l - 1].in.N,
                                   (int64_t)L.in.N, j->lv[l - 1].kappa, L.kappa);
            items.push_back(AsmBatchItem{L.in.A, L.kappa, L.val, L.dinv, L.dinv32});
        }
    }
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_TRY(assemble_batch(ctx, st, items.data(), (int)items.size()));
    for (int k = 0; k < n; ++k) {
        MgJob *j = jobs[k];
        for (int l = 0; l < j->nlev; ++l) {
            MgLevel &L = j->lv[l];
            hipLaunchKernelGGL(k_mg_val32, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.in.A->rowptr, L.val,
                               L.in.A->fixed, L.val32, L.n);
            GRMP_HIP(ctx, hipMemsetAsync(L.lam, 0, 8, st));
            hipLaunchKernelGGL(k_mg_gersh, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.in.A->rowptr, L.val, L.dinv,
                               L.in.A->fixed, L.n, (unsigned long long *)L.lam);
        }
        const MgLevel &C = j->lv[j->nlev - 1];
        hipLaunchKernelGGL(k_mg_dense_inv, dim3(1), dim3(256), 0, st, C.in.A->rowptr, C.in.A->cols, C.val,
                           C.in.A->fixed, (int)C.n, j->inv);
        GRMP_HIP(ctx, hipMemcpyAsync(j->ddesc, &j->hdesc, sizeof(MgBDesc), hipMemcpyHostToDevice, st));
        GRMP_TRY(mg_job_start(j, rtol[k], maxit));
    }
    return GRMP_OK;
}

static int mg_iterate_grid(MgJob *j, int iters);

int mg_job_iterate(MgJob *j, int iters)
{
    if (j->first_small == 0) {  // every level fits one workgroup: whole iterations in one launch
        hipLaunchKernelGGL(k_mg_pcg_block, dim3(1), dim3(MG_BT), 0, j->st, j->ddesc, j->sc, iters);
        GRMP_HIP(j->ctx, hipGetLastError());
        return GRMP_OK;
    }
    if (!j->use_graph) return mg_iterate_grid(j, iters);
    if (!j->gexec || j->graph_iters != iters) {
        if (j->gexec) (void)hipGraphExecDestroy(j->gexec);
        if (j->graph) (void)hipGraphDestroy(j->graph);
        j->gexec = nullptr;
        j->graph = nullptr;
        mg_job_collect(j);  // profiled launches made before the capture
        for (hipEvent_t e : j->pev) j->pool.push_back(e);
        j->pev.clear();
        j->pbytes.clear();
        GRMP_HIP(j->ctx, hipStreamBeginCapture(j->st, hipStreamCaptureModeThreadLocal));
        const int rc = mg_iterate_grid(j, iters);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(j->st, &g);
        GRMP_TRY(rc);
        GRMP_HIP(j->ctx, e);
        j->graph = g;
        GRMP_HIP(j->ctx, hipGraphInstantiate(&j->gexec, j->graph, nullptr, nullptr, 0));
        j->graph_iters = iters;
    }
    GRMP_HIP(j->ctx, hipGraphLaunch(j->gexec, j->st));
    return GRMP_OK;
}

// the grid-path iterations as individual launches (no memset or host sync:
// the sequence can be captured)
static int mg_iterate_grid(MgJob *j, int iters)
{
    const MgLevel &L0 = j->lv[0];
    const int64_t n = L0.n;
    for (int k = 0; k < iters; ++k) {
        MgRows a;
        rows_args(a, L0, j);
        a.v = j->p;
        a.o1 = j->q;
        a.part = j->part;
        launch_rows<OP_PCG_Q>(j, a);
        hipLaunchKernelGGL(k_mg_red<RED_ALPHA>, dim3(1), dim3(KB), 0, j->st, j->part, j->nb0, nullptr, j->sc);
        hipLaunchKernelGGL(k_mg_axpy, dim3(j->nb0), dim3(KB), 0, j->st, j->x, j->r, j->p, j->q, n, j->part, j->sc);
        hipLaunchKernelGGL(k_mg_red<RED_CONV>, dim3(1), dim3(KB), 0, j->st, j->part, j->nb0, nullptr, j->sc);
        mg_vcycle(j);
        hipLaunchKernelGGL(k_mg_red<RED_BETA>, dim3(1), dim3(KB), 0, j->st, j->part, j->nb0, nullptr, j->sc);
        hipLaunchKernelGGL(k_mg_update_p, dim3(nblk(n, 256)), dim3(256), 0, j->st, j->p, j->z, n, j->sc);
    }
    GRMP_HIP(j->ctx, hipGetLastError());
    return GRMP_OK;
}

// The updated residual met rtol: replace it by the true residual b - A x;
// accept if that meets rtol too, else restart CG from it (p = z), at most
// three times and only while each check halves the previous one.
bool mg_job_block(const MgJob *j) { return j->first_small == 0; }

// moves a set-up job to another stream (the caller has ordered it after the
// job's work on the old one); its later launches and its release go there
void mg_job_set_stream(MgJob *j, hipStream_t st) { j->st = st; }

int mg_job_verify(MgJob *j)
{
    if (j->first_small == 0) return GRMP_OK;  // k_mg_pcg_block checks in place
    // with a captured chunk, event pairs are kept for replays: the check's
    // launches stay out of the profiled set then
    struct Unprofiled {
        MgJob *j;
        bool was;
        ~Unprofiled() { j->profile = was; }
    } guard{j, j->profile};
    if (j->gexec) j->profile = false;
    const MgLevel &L0 = j->lv[0];
    const int64_t n = L0.n;
    hipLaunchKernelGGL(k_mg_verify_begin, dim3(1), dim3(1), 0, j->st, j->sc);
    MgRows a;
    rows_args(a, L0, j);
    a.v = j->x;
    a.b = j->rhs;
    a.o1 = j->r;
    launch_rows<OP_TRUE_RESID>(j, a, L0.nnz);
    hipLaunchKernelGGL(k_mg_dot, dim3(j->nb0), dim3(KB), 0, j->st, j->r, j->r, n, j->part, j->sc);
    hipLaunchKernelGGL(k_mg_red<RED_VERIFY>, dim3(1), dim3(KB), 0, j->st, j->part, j->nb0, nullptr, j->sc);
    mg_vcycle(j);
    hipLaunchKernelGGL(k_mg_red<RED_RESTART>, dim3(1), dim3(KB), 0, j->st, j->part, j->nb0, nullptr, j->sc);
    hipLaunchKernelGGL(k_mg_update_p, dim3(nblk(n, 256)), dim3(256), 0, j->st, j->p, j->z, n, j->sc);
    GRMP_HIP(j->ctx, hipGetLastError());
    return GRMP_OK;
}

int mg_job_status_async(MgJob *j, MgStatus *host)
{
    static_assert(sizeof(MgStatus) == 24, "MgStatus layout");
    MgStatus *dst = host;
    // relres, then it / done / status as laid out in MgScal
    GRMP_HIP(j->ctx, hipMemcpyAsync(&dst->relres, &j->sc->relres, 8, hipMemcpyDeviceToHost, j->st));
    GRMP_HIP(j->ctx, hipMemcpyAsync(&dst->it, &j->sc->it, 12, hipMemcpyDeviceToHost, j->st));
    GRMP_HIP(j->ctx, hipEventRecord(j->ev, j->st));
    return GRMP_OK;
}

bool mg_job_status_ready(MgJob *j) { return hipEventQuery(j->ev) != hipErrorNotReady; }

int mg_job_wait(MgJob *j)
{
    GRMP_HIP(j->ctx, hipEventSynchronize(j->ev));
    return GRMP_OK;
}

// Adds the timings of the profiled launches whose work has completed (the
// caller saw the status event of the chunk): a replayed graph keeps its event
// pairs for the next replay, individually launched ones are released.
void mg_job_collect(MgJob *j)
{
    if (!j->profile) return;
    for (size_t k = 0; k + 1 < j->pev.size(); k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, j->pev[k], j->pev[k + 1]) != hipSuccess) continue;
        j->prof_s += ms * 1e-3;
        j->prof_bytes += j->pbytes[k / 2];
        j->prof_launches += 1;
    }
    if (j->gexec) return;
    for (hipEvent_t e : j->pev) j->pool.push_back(e);
    j->pev.clear();
    j->pbytes.clear();
}

void mg_job_profile(MgJob *j, double *seconds, double *bytes, int64_t *launches)
{
    *seconds = j->prof_s;
    *bytes = j->prof_bytes;
    *launches = j->prof_launches;
}

void mg_job_destroy(MgJob *j)
{
    if (!j) return;
    if (j->gexec) (void)hipGraphExecDestroy(j->gexec);
    if (j->graph) (void)hipGraphDestroy(j->graph);
    for (hipEvent_t e : j->pev) (void)hipEventDestroy(e);
    for (hipEvent_t e : j->pool) (void)hipEventDestroy(e);
    if (j->ev) (void)hipEventDestroy(j->ev);
    if (j->buf) (void)hipFreeAsync(j->buf, j->st);
    delete j;
}

// ---------------------------------------------------- lockstep batch (host)
struct MgSeg {
    int64_t off = 0, cnt = 0;
};

struct MgBatch {
    grmp_ctx *ctx = nullptr;
    hipStream_t st = nullptr;
    bool profile = false;
    // device tables and their pinned staging (rewritten each run, after the previous run's wait)
    MgTile *dtiles = nullptr, *htiles = nullptr;
    size_t tcap = 0;
    const MgBDesc **ddesc = nullptr, **hdesc = nullptr;
    int32_t *dreq = nullptr, *hreq = nullptr;
    MgStatus *dstat = nullptr, *hstat = nullptr;
    int rcap = 0;
    hipEvent_t ev = nullptr;
    // profile: event pairs around each batched residual launch, with its bytes
    std::vector<hipEvent_t> pev, pool;  // profiled launches in flight; free events
    std::vector<double> pbytes;
    double prof_s = 0.0, prof_bytes = 0.0;
    int64_t prof_launches = 0;
};

int mg_batch_create(grmp_ctx *ctx, hipStream_t st, bool profile, MgBatch **out)
{
    MgBatch *b = new MgBatch();
    b->ctx = ctx;
    b->st = st;
    b->profile = profile;
    if (hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return grmp_fail(ctx, GRMP_ERR_HIP, "multigrid batch: event");
    }
    *out = b;
    return GRMP_OK;
}

static int batch_reserve(MgBatch *b, size_t tiles, int reqs)
{
    grmp_ctx *ctx = b->ctx;
    if (tiles > b->tcap) {
        const size_t cap = std::max(tiles, b->tcap * 2);
        if (b->dtiles) (void)hipFreeAsync(b->dtiles, b->st);
        if (b->htiles) {
            (void)hipStreamSynchronize(b->st);
            (void)hipHostFree(b->htiles);
        }
        b->dtiles = nullptr;
        b->htiles = nullptr;
        GRMP_HIP(ctx, hipMallocAsync((void **)&b->dtiles, cap * sizeof(MgTile), b->st));
        GRMP_HIP(ctx, hipHostMalloc((void **)&b->htiles, cap * sizeof(MgTile), hipHostMallocDefault));
        b->tcap = cap;
    }
    if (reqs > b->rcap) {
        const int cap = std::max(reqs, 2 * b->rcap);
        (void)hipStreamSynchronize(b->st);
        if (b->ddesc) (void)hipFree((void *)b->ddesc);
        if (b->dreq) (void)hipFree(b->dreq);
        if (b->dstat) (void)hipFree(b->dstat);
        if (b->hdesc) (void)hipHostFree((void *)b->hdesc);
        if (b->hreq) (void)hipHostFree(b->hreq);
        if (b->hstat) (void)hipHostFree(b->hstat);
        GRMP_HIP(ctx, hipMalloc((void **)&b->ddesc, cap * sizeof(MgBDesc *)));
        GRMP_HIP(ctx, hipMalloc((void **)&b->dreq, cap * sizeof(int32_t)));
        GRMP_HIP(ctx, hipMalloc((void **)&b->dstat, cap * sizeof(MgStatus)));
        GRMP_HIP(ctx, hipHostMalloc((void **)&b->hdesc, cap * sizeof(MgBDesc *), hipHostMallocDefault));
        GRMP_HIP(ctx, hipHostMalloc((void **)&b->hreq, cap * sizeof(int32_t), hipHostMallocDefault));
        GRMP_HIP(ctx, hipHostMalloc((void **)&b->hstat, cap * sizeof(MgStatus), hipHostMallocDefault));
        b->rcap = cap;
    }
    return GRMP_OK;
}

static inline unsigned ublk(int64_t n) { return (unsigned)n; }

// Enqueues iters lockstep iterations of the n jobs (all set up on the
// batch's stream), then a status copy for each job into mg_batch_status().
int mg_batch_run(MgBatch *b, MgJob *const *jobs, int n, int iters)
{
    grmp_ctx *ctx = b->ctx;
    hipStream_t st = b->st;
    std::vector<int> grid, blk;
    int Fmax = 0;
    for (int i = 0; i < n; ++i) {
        if (jobs[i]->first_small == 0) blk.push_back(i);
        else {
            grid.push_back(i);
            Fmax = std::max(Fmax, jobs[i]->first_small);
        }
    }
    // tile segments: level-0 rows; per grid level l: its rows, the next
    // level's rows (restriction), its prolongation items
    std::vector<MgTile> T;
    auto add = [&](MgSeg &sg, int lev, bool next, bool items) {
        sg.off = (int64_t)T.size();
        for (int g : grid) {
            const MgJob *j = jobs[g];
            if (j->first_small <= lev) continue;
            const int L = next ? lev + 1 : lev;
            int64_t cnt = j->lv[L].n;
            if (items && j->lv[lev].in.fe == GRMP_FE_H1P2) {
                const int64_t N = j->lv[lev].in.N;
                cnt = std::max((N + 1) * (N + 1), 2 * N * N);
            }
            for (int64_t r0 = 0; r0 < cnt; r0 += (int64_t)KB * MG_TILE_CHUNKS) T.push_back(MgTile{g, L, (int32_t)r0, 0});
        }
        sg.cnt = (int64_t)T.size() - sg.off;
    };
    std::vector<MgSeg> lev(Fmax), res(Fmax), pro(Fmax);
    for (int l = 0; l < Fmax; ++l) {
        add(lev[l], l, false, false);
        add(res[l], l, true, false);
        add(pro[l], l, false, true);
    }
    GRMP_TRY(batch_reserve(b, std::max<size_t>(T.size(), 1), std::max(n, 1)));
    memcpy(b->htiles, T.data(), T.size() * sizeof(MgTile));
    for (int i = 0; i < n; ++i) b->hdesc[i] = jobs[i]->ddesc;
    for (size_t k = 0; k < grid.size(); ++k) b->hreq[k] = grid[k];
    for (size_t k = 0; k < blk.size(); ++k) b->hreq[grid.size() + k] = blk[k];
    if (!T.empty())
        GRMP_HIP(ctx, hipMemcpyAsync(b->dtiles, b->htiles, T.size() * sizeof(MgTile), hipMemcpyHostToDevice, st));
    GRMP_HIP(ctx, hipMemcpyAsync((void *)b->ddesc, (const void *)b->hdesc, n * sizeof(MgBDesc *), hipMemcpyHostToDevice,
                                 st));
    GRMP_HIP(ctx, hipMemcpyAsync(b->dreq, b->hreq, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    const MgBDesc *const *D = b->ddesc;
    const int32_t *G = b->dreq, *Bk = b->dreq + grid.size();
    const int ng = (int)grid.size(), nbk = (int)blk.size();
    auto tiles = [&](const MgSeg &sg) { return b->dtiles + sg.off; };
    for (int k = 0; k < iters && ng > 0; ++k) {
        const MgSeg &L0 = lev[0];
        hipLaunchKernelGGL(k_mgb_rows<OP_PCG_Q>, dim3(ublk(L0.cnt)), dim3(KB), 0, st, tiles(L0), D);
        hipLaunchKernelGGL(k_mgb_red<RED_ALPHA>, dim3(ng), dim3(KB), 0, st, D, G);
        hipLaunchKernelGGL(k_mgb_axpy, dim3(ublk(L0.cnt)), dim3(KB), 0, st, tiles(L0), D);
        hipLaunchKernelGGL(k_mgb_red<RED_CONV>, dim3(ng), dim3(KB), 0, st, D, G);
        for (int l = 0; l < Fmax; ++l) {
            if (lev[l].cnt == 0) continue;
            hipLaunchKernelGGL(k_mgb_first, dim3(ublk(lev[l].cnt)), dim3(KB), 0, st, tiles(lev[l]), D);
            hipLaunchKernelGGL(k_mgb_rows<OP_CHEB_PRE>, dim3(ublk(lev[l].cnt)), dim3(KB), 0, st, tiles(lev[l]), D);
            hipEvent_t e0, e1;
            if (b->profile && ev_take(b->pool, e0, e1)) {
                hipExtLaunchKernelGGL(k_mgb_rows<OP_RESID>, dim3(ublk(lev[l].cnt)), dim3(KB), 0, st, e0, e1, 0,
                                      tiles(lev[l]), D);
                b->pev.push_back(e0);
                b->pev.push_back(e1);
                double bytes = 0.0;
                for (int g : grid)
                    if (jobs[g]->first_small > l) {
                        const MgLevel &Lv = jobs[g]->lv[l];
                        bytes += (double)Lv.nnz * 8.0 + 4.0 * (Lv.n + 1) + 25.0 * Lv.n;
                    }
                b->pbytes.push_back(bytes);
            } else {
                hipLaunchKernelGGL(k_mgb_rows<OP_RESID>, dim3(ublk(lev[l].cnt)), dim3(KB), 0, st, tiles(lev[l]), D);
            }
            hipLaunchKernelGGL(k_mgb_restrict, dim3(ublk(res[l].cnt)), dim3(KB), 0, st, tiles(res[l]), D);
        }
        hipLaunchKernelGGL(k_mgb_vcycle_block, dim3(ng), dim3(MG_BT), 0, st, D, G);
        for (int l = Fmax - 1; l >= 0; --l) {
            if (lev[l].cnt == 0) continue;
            hipLaunchKernelGGL(k_mgb_prolong, dim3(ublk(pro[l].cnt)), dim3(KB), 0, st, tiles(pro[l]), D);
            hipLaunchKernelGGL(k_mgb_rows<OP_CHEB_POST1>, dim3(ublk(lev[l].cnt)), dim3(KB), 0, st, tiles(lev[l]), D);
            hipLaunchKernelGGL(k_mgb_rows<OP_CHEB_POST2>, dim3(ublk(lev[l].cnt)), dim3(KB), 0, st, tiles(lev[l]), D);
        }
        hipLaunchKernelGGL(k_mgb_red<RED_BETA>, dim3(ng), dim3(KB), 0, st, D, G);
        hipLaunchKernelGGL(k_mgb_update_p, dim3(ublk(L0.cnt)), dim3(KB), 0, st, tiles(L0), D);
    }
    if (nbk > 0) hipLaunchKernelGGL(k_mgb_pcg_block, dim3(nbk), dim3(MG_BT), 0, st, D, Bk, iters);
    hipLaunchKernelGGL(k_mgb_status, dim3(nblk(n, 256)), dim3(256), 0, st, D, n, b->dstat);
    GRMP_HIP(ctx, hipMemcpyAsync(b->hstat, b->dstat, n * sizeof(MgStatus), hipMemcpyDeviceToHost, st));
    GRMP_HIP(ctx, hipEventRecord(b->ev, st));
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

int mg_batch_wait(MgBatch *b)
{
    GRMP_HIP(b->ctx, hipEventSynchronize(b->ev));
    for (size_t k = 0; k + 1 < b->pev.size(); k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->pev[k], b->pev[k + 1]) == hipSuccess) {
            b->prof_s += ms * 1e-3;
            b->prof_bytes += b->pbytes[k / 2];
            b->prof_launches += 1;
        }
        b->pool.push_back(b->pev[k]);
        b->pool.push_back(b->pev[k + 1]);
    }
    b->pev.clear();
    b->pbytes.clear();
    return GRMP_OK;
}

const MgStatus *mg_batch_status(const MgBatch *b) { return b->hstat; }

void mg_batch_profile(const MgBatch *b, double *seconds, double *bytes, int64_t *launches)
{
    *seconds = b->prof_s;
    *bytes = b->prof_bytes;
    *launches = b->prof_launches;
}

void mg_batch_destroy(MgBatch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->st);
    if (b->dtiles) (void)hipFree(b->dtiles);
    if (b->htiles) (void)hipHostFree(b->htiles);
    if (b->ddesc) (void)hipFree((void *)b->ddesc);
    if (b->dreq) (void)hipFree(b->dreq);
    if (b->dstat) (void)hipFree(b->dstat);
    if (b->hdesc) (void)hipHostFree((void *)b->hdesc);
    if (b->hreq) (void)hipHostFree(b->hreq);
    if (b->hstat) (void)hipHostFree(b->hstat);
    for (hipEvent_t e : b->pev) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->pool) (void)hipEventDestroy(e);
    if (b->ev) (void)hipEventDestroy(b->ev);
    delete b;
}

// ------------------------------------------------- one multigrid solve
// The served path's V-cycle CG for one caller-built hierarchy: level 0 is
// the caller's system (its forms are assembled by the job with the job's
// own cell coefficient), coarser levels are rediscretizations on the unit
// squares of ceil(N/2) down to N <= 4.  The solution is 0 on fixed rows
// (homogeneous Dirichlet data).  A handle keeps the assembled hierarchy for
// further right-hand sides, as the reference keeps its factorization while
// the matrix is unchanged.  Iterations are chunked as in the server: status
// after each chunk, the true-residual check when the scalars ask.
struct grmp_mg {
    grmp_ctx *ctx;
    MgJob *job;
    MgStatus *h;  // pinned
    const grmp_matrix *A0;
    int64_t n;
    // a second copy of the hierarchy's values and vectors on a stream of its
    // own, so that two components' V-cycles run at the same time
    // (mg_precond_apply_pair); built on first use
    std::vector<MgLevelIn> in;
    MgJob *job2 = nullptr;
    hipStream_t st2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

extern "C" int grmp_mg_destroy(grmp_mg *m)
{
    if (!m) return GRMP_OK;
    if (m->job) {
        (void)hipStreamSynchronize(m->ctx->stream);
        mg_job_destroy(m->job);
        (void)hipStreamSynchronize(m->ctx->stream);
    }
    if (m->st2) {
        (void)hipStreamSynchronize(m->st2);
        if (m->job2) mg_job_destroy(m->job2);
        (void)hipStreamSynchronize(m->st2);
        (void)hipStreamDestroy(m->st2);
    }
    if (m->fork) (void)hipEventDestroy(m->fork);
    if (m->join) (void)hipEventDestroy(m->join);
    if (m->h) (void)hipHostFree(m->h);
    delete m;
    return GRMP_OK;
}

extern "C" int grmp_mg_create(grmp_ctx *ctx, int nlev, const grmp_matrix *const *levels,
                              const grmp_space *const *spaces, const grmp_vector *kappa0, grmp_mg **out)
{
    if (!ctx || !levels || !spaces || !kappa0 || !out) return GRMP_ERR_ARG;
    *out = nullptr;
    GRMP_CHECK(ctx, nlev >= 1 && nlev <= MG_MAX_LEVELS, GRMP_ERR_ARG, "grmp_mg_create: %d levels", nlev);
    for (int l = 0; l < nlev; ++l)
        GRMP_CHECK(ctx, levels[l] && spaces[l] && levels[l]->ctx == ctx && levels[l]->nblocks == 1 &&
                            levels[l]->spaces[0] == spaces[l],
                   GRMP_ERR_ARG, "grmp_mg_create: level %d is not a one-block matrix over its space", l);
    const grmp_mesh *m0 = spaces[0]->mesh;
    GRMP_CHECK(ctx, kappa0->ctx == ctx && kappa0->n == m0->ncells, GRMP_ERR_ARG,
               "grmp_mg_create: kappa0 needs one value per level-0 cell (%lld)", (long long)m0->ncells);
    std::vector<MgLevelIn> in(nlev);
    for (int l = 0; l < nlev; ++l)
        in[l] = MgLevelIn{levels[l], spaces[l], spaces[l]->mesh->N, spaces[l]->fe, spaces[l]->ncomp};
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = ctx->stream;
    grmp_mg *m = new grmp_mg{ctx, nullptr, nullptr, levels[0], levels[0]->n};
    m->in = in;
    int rc = GRMP_OK;
    do {
        if ((rc = mg_job_create(ctx, st, in.data(), nlev, false, 2048, 2048, false, &m->job))) break;
        if (hipHostMalloc((void **)&m->h, sizeof(MgStatus), hipHostMallocDefault) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mg_create: status buffer");
            break;
        }
        if (hipMemcpyAsync(mg_job_kappa0(m->job), kappa0->d, sizeof(double) * m0->ncells, hipMemcpyDeviceToDevice,
                           st) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mg_create: coefficient");
            break;
        }
        if ((rc = mg_job_assemble(m->job))) break;
        if (hipStreamSynchronize(st) != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mg_create: assembly");
    } while (0);
    if (rc) {
        grmp_mg_destroy(m);
        return rc;
    }
    *out = m;
    return GRMP_OK;
}

extern "C" int grmp_mg_apply(grmp_mg *m, const grmp_vector *b, grmp_vector *x, double rtol, int64_t maxit,
                             grmp_solve_info *info)
{
    if (!m || !b || !x) return GRMP_ERR_ARG;
    grmp_ctx *ctx = m->ctx;
    GRMP_CHECK(ctx, b->ctx == ctx && x->ctx == ctx && b->n == m->n && x->n == m->n && b->d != x->d, GRMP_ERR_ARG,
               "grmp_mg_apply: b and x need the level-0 length %lld and distinct storage", (long long)m->n);
    GRMP_CHECK(ctx, rtol > 0.0 && rtol < 1.0 && maxit >= 1 && maxit <= (1 << 30), GRMP_ERR_ARG,
               "grmp_mg_apply: rtol in (0, 1) and maxit >= 1 needed");
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = ctx->stream;
    const double t0 = wall_seconds();
    MgJob *j = m->job;
    MgStatus *h = m->h;
    memset(h, 0, sizeof(MgStatus));
    int rc = GRMP_OK;
    do {
        if (hipMemcpyAsync(mg_job_rhs(j), b->d, sizeof(double) * m->n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mg_apply: right-hand side");
            break;
        }
        if ((rc = zero_fixed_launch(m->A0, mg_job_rhs(j), st))) break;
        if ((rc = mg_job_start(j, rtol, (int)maxit))) break;
        if ((rc = mg_job_iterate(j, 4)) || (rc = mg_job_status_async(j, h)) || (rc = mg_job_wait(j))) break;
        while (h->done != 1) {
            if (h->done == 2) {
                rc = mg_job_verify(j);
                if (!rc && mg_job_block(j)) rc = mg_job_iterate(j, 1);
            } else {
                rc = mg_job_iterate(j, 4);
            }
            if (rc || (rc = mg_job_status_async(j, h)) || (rc = mg_job_wait(j))) break;
        }
        if (rc) break;
        if (hipMemcpyAsync(x->d, mg_job_solution(j), sizeof(double) * m->n, hipMemcpyDeviceToDevice, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mg_apply: solution");
    } while (0);
    grmp_solve_info inf;
    memset(&inf, 0, sizeof(inf));
    inf.status = h->status;
    inf.iterations = h->it;
    inf.relres = h->relres;
    inf.outer = 1;
    inf.seconds = wall_seconds() - t0;
    if (info) *info = inf;
    if (rc) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    if (inf.status == GRMP_ERR_NOCONV)
        return grmp_fail(ctx, GRMP_ERR_NOCONV, "grmp_mg_apply: no convergence, relres %.3e after %lld iterations",
                         inf.relres, (long long)inf.iterations);
    return inf.status;
}

// ------------------------------------------- one V-cycle as a preconditioner
// r_0 = v with the fixed rows zeroed; the device scalars are reset so the
// V-cycle kernels run (a finished CG leaves done set)
__global__ void k_mg_pc_in(const double *__restrict__ v, const uint8_t *__restrict__ fixed, double *__restrict__ r,
                           MgScal *__restrict__ sc, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        sc->done = 0;
        sc->status = 0;
    }
    if (i < n) r[i] = fixed[i] ? 0.0 : v[i];
}

int64_t mg_level0_rows(const grmp_mg *m) { return m->n; }

// z = V v: one V-cycle of the hierarchy from a zero start on stream st (the
// hierarchy's own stream): symmetric and positive definite, so it can
// precondition MINRES as it does CG.  z and v are n-vectors of the level-0
// numbering; fixed rows of z are 0.
// the second hierarchy of mg_precond_apply_pair: the same levels assembled
// from the same coefficient, so its V-cycle gives the first one's bits
int mg_pair_prepare(grmp_mg *m)
{
    if (m->job2) return GRMP_OK;
    grmp_ctx *ctx = m->ctx;
    if (!m->st2 && hipStreamCreateWithFlags(&m->st2, hipStreamNonBlocking) != hipSuccess) {
        m->st2 = nullptr;
        return grmp_fail(ctx, GRMP_ERR_HIP, "multigrid pair: stream");
    }
    if ((!m->fork && hipEventCreateWithFlags(&m->fork, hipEventDisableTiming) != hipSuccess) ||
        (!m->join && hipEventCreateWithFlags(&m->join, hipEventDisableTiming) != hipSuccess))
        return grmp_fail(ctx, GRMP_ERR_HIP, "multigrid pair: events");
    MgJob *j2 = nullptr;
    GRMP_TRY(mg_job_create(ctx, m->st2, m->in.data(), (int)m->in.size(), false, 2048, 2048, false, &j2));
    int rc = GRMP_OK;
    // the first hierarchy's coefficient is complete on the context stream
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpyAsync(mg_job_kappa0(j2), mg_job_kappa0(m->job), sizeof(double) * m->in[0].sp->mesh->ncells,
                       hipMemcpyDeviceToDevice, m->st2) != hipSuccess)
        rc = grmp_fail(ctx, GRMP_ERR_HIP, "multigrid pair: coefficient");
    if (!rc) rc = mg_job_assemble(j2);
    if (!rc && hipStreamSynchronize(m->st2) != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "multigrid pair: assembly");
    if (rc) {
        (void)hipStreamSynchronize(m->st2);
        mg_job_destroy(j2);
        return rc;
    }
    m->job2 = j2;
    return GRMP_OK;
}

// z0 = V v0 on the hierarchy's stream st and z1 = V v1 on the second
// hierarchy's stream, at the same time; st waits for both
int mg_precond_apply_pair(grmp_mg *m, const double *v0, double *z0, const double *v1, double *z1, hipStream_t st)
{
    MgJob *j = m->job, *j2 = m->job2;
    const int64_t n = m->n;
    GRMP_CHECK(m->ctx, st == j->st && j2, GRMP_ERR_ARG, "multigrid preconditioner pair: not prepared");
    GRMP_HIP(m->ctx, hipEventRecord(m->fork, st));
    GRMP_HIP(m->ctx, hipStreamWaitEvent(m->st2, m->fork, 0));
    hipLaunchKernelGGL(k_mg_pc_in, dim3(nblk(n, 256)), dim3(256), 0, m->st2, v1, m->A0->fixed, j2->r, j2->sc, n);
    mg_vcycle(j2);
    GRMP_HIP(m->ctx, hipMemcpyAsync(z1, j2->z, sizeof(double) * n, hipMemcpyDeviceToDevice, m->st2));
    hipLaunchKernelGGL(k_mg_pc_in, dim3(nblk(n, 256)), dim3(256), 0, st, v0, m->A0->fixed, j->r, j->sc, n);
    mg_vcycle(j);
    GRMP_HIP(m->ctx, hipMemcpyAsync(z0, j->z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    GRMP_HIP(m->ctx, hipEventRecord(m->join, m->st2));
    GRMP_HIP(m->ctx, hipStreamWaitEvent(st, m->join, 0));
    GRMP_HIP(m->ctx, hipGetLastError());
    return GRMP_OK;
}

int mg_precond_apply(grmp_mg *m, const double *v, double *z, hipStream_t st)
{
    MgJob *j = m->job;
    const int64_t n = m->n;
    GRMP_CHECK(m->ctx, st == j->st, GRMP_ERR_ARG, "multigrid preconditioner: not the hierarchy's stream");
    hipLaunchKernelGGL(k_mg_pc_in, dim3(nblk(n, 256)), dim3(256), 0, st, v, m->A0->fixed, j->r, j->sc, n);
    mg_vcycle(j);
    GRMP_HIP(m->ctx, hipMemcpyAsync(z, j->z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    GRMP_HIP(m->ctx, hipGetLastError());
    return GRMP_OK;
}

extern "C" int grmp_mg_solve(grmp_ctx *ctx, int nlev, const grmp_matrix *const *levels,
                             const grmp_space *const *spaces, const grmp_vector *kappa0, const grmp_vector *b,
                             grmp_vector *x, double rtol, int64_t maxit, grmp_solve_info *info)
{
    grmp_mg *m = nullptr;
    GRMP_TRY(grmp_mg_create(ctx, nlev, levels, spaces, kappa0, &m));
    const int rc = grmp_mg_apply(m, b, x, rtol, maxit, info);
    grmp_mg_destroy(m);
    return rc;
}

// ================================================================ AMG
// Algebraic multigrid for symmetric positive definite systems on any mesh
// (the geometric hierarchy above needs generated unit squares).
//   aggregation   two passes of pairwise matching per level: a free row pairs
//                 with its strongest free neighbour (the largest positive
//                 -a_ij / sqrt(a_ii a_jj), ties to the smaller index) when the
//                 choice is mutual, in three handshake rounds; a row left over
//                 joins the pair of its strongest matched neighbour, or stays
//                 single when it has none.
//                 No absolute threshold: on coarse 3D levels with 20-30
//                 neighbours per row every normalized coupling is small.
//   coarse level  the Galerkin product P^T A P of the piecewise-constant P:
//                 every entry keyed (agg i, agg j), stably sorted, runs summed.
//   cycle         V(1,1) with the Chebyshev smoothing of the geometric levels
//                 (FP32-stored values, FP64 vectors), the coarse correction
//                 scaled by AMG_SCALE (plain aggregation under-corrects smooth
//                 error; the scaled cycle stays symmetric and positive
//                 definite for a scale below 2), the dense inverse on the
//                 coarsest level (at most MG_DENSE_MAX rows).
//   Krylov        FP64 CG with the device scalars and the true-residual check
//                 of the geometric solver.
// Fixed (penalized) rows are eliminated: x_d = b_d / A_dd, and they belong to
// no aggregate.  Every step is deterministic: handshakes read the previous
// round's state only, the sorts are stable and the sums run in a fixed order.
#define AMG_MAX_LEVELS 24
#define AMG_THETA 0.0
// 1.8 measured best over 1.3-1.9: 48 / 66 / 35 iterations on the 2048^2 P1,
// 2048^2 P2 and 160^3 tet Poisson systems, against 70 / 101 / 36 with 1.5
#define AMG_SCALE 1.8

struct AmgLevel {
    int64_t n = 0, nnz = 0;
    const int32_t *rp = nullptr, *ci = nullptr;  // level 0: the matrix's pattern
    const double *val = nullptr;                 // level 0: the matrix's FP64 values
    const uint8_t *fixed = nullptr;              // coarse levels: all zero
    float *val32 = nullptr;
    double *dinv = nullptr, *lam = nullptr;
    double *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;  // level 0: x = z, b = r of the CG
    int32_t *agg = nullptr;                      // row -> row of the next level (-1: fixed)
    int32_t *mptr = nullptr, *midx = nullptr;    // rows of this level: their members on the previous one
};

struct grmp_amg {
    grmp_ctx *ctx = nullptr;
    const grmp_matrix *A = nullptr;
    int nlev = 0;
    AmgLevel lv[AMG_MAX_LEVELS];
    double *inv = nullptr;
    double *x = nullptr, *rhs = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double *part = nullptr, *part2 = nullptr;
    MgScal *sc = nullptr;
    int nb0 = 0;
    MgStatus *h = nullptr;  // pinned
    hipEvent_t ev = nullptr;
    std::vector<void *> mem;  // device allocations of the hierarchy
    int64_t nnz_total = 0;    // summed nonzeros of every level
};

template <class T>
static int amg_alloc(grmp_amg *m, T **p, size_t count)
{
    GRMP_TRY(dalloc(m->ctx, p, count));
    m->mem.push_back((void *)*p);
    return GRMP_OK;
}

// diagonal of each row (0 if it has none)
__global__ void k_amg_diag(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ val, int64_t n, double *__restrict__ diag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) d = val[k];
    diag[i] = d;
}

// -(a_ij + a_ji) / 2: the strength of the pair (i, j) seen from both rows, so
// that both rows of a pair rank it alike when A is nonsymmetric (a Newton
// matrix); equal to -a_ij bit for bit when a_ji == a_ij.  Columns are
// ascending within a row (pattern and Galerkin levels alike); a_ji missing
// from row j counts as a_ij.
__device__ __forceinline__ double amg_pair_coupling(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ val, int32_t i, int32_t j, double aij)
{
    int32_t lo = rp[j], hi = rp[j + 1];
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (ci[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    const double aji = (lo < rp[j + 1] && ci[lo] == i) ? val[lo] : aij;
    return -0.5 * (aij + aji);
}

// A key of the pair {i, j}, the same from both rows: the SplitMix64 finalizer
// (Steele, Lea, Flood 2014) of the ordered pair.  Equal strengths, which a
// structured mesh gives everywhere, are ranked by it; ranking them by the
// neighbour's index lets choices chain instead of pairing, and the coarsening
// of such a level stalls (P2 on 256 x 256 squares: 65029 -> 62672 rows).
__device__ __forceinline__ uint64_t amg_pair_key(int32_t i, int32_t j)
{
    const uint64_t lo = (uint64_t)(uint32_t)(i < j ? i : j), hi = (uint64_t)(uint32_t)(i < j ? j : i);
    uint64_t z = (lo << 32 | hi) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// best[i]: the strongest free, unmatched neighbour of the free, unmatched row
// i, or -1 (fixed == nullptr: no fixed rows); ties go to the larger pair key,
// then to the smaller index, so that strongest pairs are chosen from both ends
__global__ void k_amg_best(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ val, const double *__restrict__ diag,
                           const uint8_t *__restrict__ fixed, const int32_t *__restrict__ mate, int64_t n,
                           double theta, int32_t *__restrict__ best)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t bj = -1;
    if (mate[i] < 0 && !(fixed && fixed[i]) && diag[i] > 0.0) {
        double bs = 0.0;
        uint64_t bk = 0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int32_t j = ci[k];
            if (j == i || mate[j] >= 0 || (fixed && fixed[j]) || !(diag[j] > 0.0)) continue;
            const double st = amg_pair_coupling(rp, ci, val, (int32_t)i, j, val[k]) / sqrt(diag[i] * diag[j]);
            if (!(st > theta)) continue;
            const uint64_t key = amg_pair_key((int32_t)i, j);
            if (bj < 0 || st > bs || (st == bs && (key > bk || (key == bk && j < bj)))) {
                bs = st;
                bk = key;
                bj = j;
            }
        }
    }
    best[i] = bj;
}

// mutual choices pair up (each row writes only its own mate)
__global__ void k_amg_match(const int32_t *__restrict__ best, int32_t *__restrict__ mate, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t j = best[i];
    if (mate[i] < 0 && j >= 0 && best[j] == (int32_t)i) mate[i] = j;
}

// a free row left unmatched joins the pair of its strongest matched
// neighbour (join[i] = that neighbour), else stays single (join[i] = -1)
__global__ void k_amg_join(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ val, const double *__restrict__ diag,
                           const uint8_t *__restrict__ fixed, const int32_t *__restrict__ mate, int64_t n,
                           int32_t *__restrict__ join)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t bj = -1;
    if (mate[i] < 0 && !(fixed && fixed[i]) && diag[i] > 0.0) {
        double bs = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int32_t j = ci[k];
            if (j == i || mate[j] < 0 || (fixed && fixed[j]) || !(diag[j] > 0.0)) continue;
            const double st = amg_pair_coupling(rp, ci, val, (int32_t)i, j, val[k]) / sqrt(diag[i] * diag[j]);
            if (!(st > 0.0)) continue;
            if (bj < 0 || st > bs || (st == bs && j < bj)) {
                bs = st;
                bj = j;
            }
        }
    }
    join[i] = bj;
}

// lead[i] = 1 for the first row of each pair and for rows left single
__global__ void k_amg_lead(const int32_t *__restrict__ mate, const int32_t *__restrict__ join,
                           const uint8_t *__restrict__ fixed, int64_t n, int32_t *__restrict__ lead)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fixed && fixed[i]) {
        lead[i] = 0;
        return;
    }
    const int32_t m = mate[i];
    lead[i] = m >= 0 ? ((int32_t)i <= m ? 1 : 0) : (join[i] < 0 ? 1 : 0);
}

__global__ void k_amg_assign(const int32_t *__restrict__ mate, const int32_t *__restrict__ join,
                             const uint8_t *__restrict__ fixed, const int32_t *__restrict__ id, int64_t n,
                             int32_t *__restrict__ agg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fixed && fixed[i]) {
        agg[i] = -1;
        return;
    }
    int32_t m = mate[i], r = (int32_t)i;
    if (m < 0 && join[i] >= 0) {  // the pair of the neighbour it joined
        r = join[i];
        m = mate[r];
    }
    agg[i] = m < 0 ? id[r] : id[m < r ? m : r];
}

// agg[i] = a2[a1[i]] (composition of two pairwise passes)
__global__ void k_amg_compose(const int32_t *__restrict__ a1, const int32_t *__restrict__ a2, int64_t n,
                              int32_t *__restrict__ agg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    agg[i] = a1[i] < 0 ? -1 : a2[a1[i]];
}

// Galerkin keys: entry (i, j) -> agg i * nc + agg j; entries of fixed rows or
// columns get the sentinel nc * nc (sorted last, dropped)
__global__ void k_amg_keys(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ val, const int32_t *__restrict__ agg, int64_t n, int64_t nc,
                           uint64_t *__restrict__ keys, double *__restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t ai = agg[i];
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int32_t aj = agg[ci[k]];
        const bool in = ai >= 0 && aj >= 0;
        keys[k] = in ? (uint64_t)ai * (uint64_t)nc + (uint64_t)aj : (uint64_t)nc * (uint64_t)nc;
        vals[k] = in ? val[k] : 0.0;
    }
}

// the coarse CSR from the sorted unique keys (rows ascending, columns ascending within a row)
__global__ void k_amg_rowcount(const uint64_t *__restrict__ ukeys, int64_t m, int64_t nc, int32_t *__restrict__ cnt)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= m) return;
    atomicAdd(&cnt[ukeys[u] / (uint64_t)nc], 1);
}

__global__ void k_amg_cols(const uint64_t *__restrict__ ukeys, int64_t m, int64_t nc, int32_t *__restrict__ cols)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= m) return;
    cols[u] = (int32_t)(ukeys[u] % (uint64_t)nc);
}

// members of each aggregate: keys agg (fixed rows -> nc, sorted last), values the row
__global__ void k_amg_member_keys(const int32_t *__restrict__ agg, int64_t n, int32_t nc, int32_t *__restrict__ key,
                                  int32_t *__restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = agg[i] < 0 ? nc : agg[i];
    idx[i] = (int32_t)i;
}

__global__ void k_amg_member_count(const int32_t *__restrict__ agg, int64_t n, int32_t *__restrict__ cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (agg[i] >= 0) atomicAdd(&cnt[agg[i]], 1);
}

__global__ void k_amg_dinv(const double *__restrict__ diag, const uint8_t *__restrict__ fixed, int64_t n,
                           double *__restrict__ dinv)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dinv[i] = (fixed[i] || !(diag[i] > 0.0)) ? 0.0 : 1.0 / diag[i];
}

// restriction b_c[I] = sum of r over the members of I, in row order
__global__ void k_amg_restrict(const int32_t *__restrict__ mptr, const int32_t *__restrict__ midx,
                               const double *__restrict__ r, double *__restrict__ bc, int64_t nc, const MgScal *sc)
{
    if (sc->done) return;
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nc) return;
    double s = 0.0;
    for (int32_t k = mptr[I]; k < mptr[I + 1]; ++k) s += r[midx[k]];
    bc[I] = s;
}

// prolongation x_i += scale x_c[agg i] on the free rows
__global__ void k_amg_prolong(const int32_t *__restrict__ agg, const double *__restrict__ xc, double scale,
                              double *__restrict__ x, int64_t n, const MgScal *sc)
{
    if (sc->done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t a = agg[i];
    if (a >= 0) x[i] += scale * xc[a];
}

// One pairwise pass over a level (rp, ci, val, fixed): agg[n], *nc.
static int amg_pairwise(grmp_ctx *ctx, const int32_t *rp, const int32_t *ci, const double *val, const uint8_t *fixed,
                        int64_t n, int32_t *agg, int64_t *nc)
{
    hipStream_t st = ctx->stream;
    double *diag = nullptr;
    int32_t *mate = nullptr, *best = nullptr, *lead = nullptr, *id = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &diag, n)) || (rc = dalloc(ctx, &mate, n)) || (rc = dalloc(ctx, &best, n)) ||
            (rc = dalloc(ctx, &lead, n)) || (rc = dalloc(ctx, &id, n + 1)))
            break;
        hipLaunchKernelGGL(k_amg_diag, dim3(nblk(n, 256)), dim3(256), 0, st, rp, ci, val, n, diag);
        if (hipMemsetAsync(mate, 0xff, sizeof(int32_t) * n, st) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "AMG aggregation: memset");
            break;
        }
        for (int round = 0; round < 3; ++round) {
            hipLaunchKernelGGL(k_amg_best, dim3(nblk(n, 256)), dim3(256), 0, st, rp, ci, val, diag, fixed, mate, n,
                               AMG_THETA, best);
            hipLaunchKernelGGL(k_amg_match, dim3(nblk(n, 256)), dim3(256), 0, st, best, mate, n);
        }
        hipLaunchKernelGGL(k_amg_join, dim3(nblk(n, 256)), dim3(256), 0, st, rp, ci, val, diag, fixed, mate, n, best);
        hipLaunchKernelGGL(k_amg_lead, dim3(nblk(n, 256)), dim3(256), 0, st, mate, best, fixed, n, lead);
        if ((rc = exclusive_scan_i32(ctx, lead, id, n))) break;
        int32_t cnt = 0;
        if (hipMemcpyAsync(&cnt, id + n, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "AMG aggregation: count");
            break;
        }
        hipLaunchKernelGGL(k_amg_assign, dim3(nblk(n, 256)), dim3(256), 0, st, mate, best, fixed, id, n, agg);
        *nc = cnt;
        if (hipGetLastError() != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "AMG aggregation: launch");
    } while (0);
    dfree(ctx, diag);
    dfree(ctx, mate);
    dfree(ctx, best);
    dfree(ctx, lead);
    dfree(ctx, id);
    return rc;
}

// Galerkin product of a level with the aggregation agg (nc aggregates): the
// coarse CSR in new device arrays (caller frees, or keeps via amg_alloc-owned copies)
static int amg_galerkin(grmp_ctx *ctx, const int32_t *rp, const int32_t *ci, const double *val, int64_t n, int64_t nnz,
                        const int32_t *agg, int64_t nc, int32_t **crp, int32_t **cci, double **cval, int64_t *cnnz)
{
    hipStream_t st = ctx->stream;
    *crp = nullptr;
    *cci = nullptr;
    *cval = nullptr;
    uint64_t *k0 = nullptr, *k1 = nullptr, *uk = nullptr;
    double *v0 = nullptr, *v1 = nullptr, *us = nullptr;
    int32_t *cnt = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &k0, nnz)) || (rc = dalloc(ctx, &k1, nnz)) || (rc = dalloc(ctx, &v0, nnz)) ||
            (rc = dalloc(ctx, &v1, nnz)) || (rc = dalloc(ctx, &uk, nnz)) || (rc = dalloc(ctx, &us, nnz)))
            break;
        hipLaunchKernelGGL(k_amg_keys, dim3(nblk(n, 256)), dim3(256), 0, st, rp, ci, val, agg, n, nc, k0, v0);
        int bits = 1;
        while (bits < 64 && ((uint64_t)1 << bits) <= (uint64_t)nc * (uint64_t)nc) ++bits;
        if ((rc = sort_pairs_u64_f64(ctx, k0, k1, v0, v1, nnz, bits))) break;
        int64_t m = 0;
        if ((rc = reduce_runs_u64_f64(ctx, k1, v1, nnz, uk, us, &m))) break;
        // the sentinel run (entries of fixed rows or columns) is the last one
        uint64_t last = 0;
        if (m > 0) {
            if (hipMemcpyAsync(&last, uk + (m - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                rc = grmp_fail(ctx, GRMP_ERR_HIP, "AMG Galerkin: readback");
                break;
            }
            if (last == (uint64_t)nc * (uint64_t)nc) --m;
        }
        if ((rc = dalloc(ctx, &cnt, nc)) || (rc = dalloc(ctx, crp, nc + 1)) || (rc = dalloc(ctx, cci, m)) ||
            (rc = dalloc(ctx, cval, m)))
            break;
        if (hipMemsetAsync(cnt, 0, sizeof(int32_t) * nc, st) != hipSuccess ||
            hipMemcpyAsync(*cval, us, sizeof(double) * m, hipMemcpyDeviceToDevice, st) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "AMG Galerkin: copy");
            break;
        }
        hipLaunchKernelGGL(k_amg_rowcount, dim3(nblk(m, 256)), dim3(256), 0, st, uk, m, nc, cnt);
        hipLaunchKernelGGL(k_amg_cols, dim3(nblk(m, 256)), dim3(256), 0, st, uk, m, nc, *cci);
        if ((rc = exclusive_scan_i32(ctx, cnt, *crp, nc))) break;
        *cnnz = m;
        if (hipGetLastError() != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "AMG Galerkin: launch");
    } while (0);
    dfree(ctx, k0);
    dfree(ctx, k1);
    dfree(ctx, v0);
    dfree(ctx, v1);
    dfree(ctx, uk);
    dfree(ctx, us);
    dfree(ctx, cnt);
    if (rc) {
        dfree(ctx, *crp);
        dfree(ctx, *cci);
        dfree(ctx, *cval);
    }
    return rc;
}

// smoother data of a level: FP32 values, inverse diagonal, Gershgorin bound
static int amg_level_smoother(grmp_amg *m, AmgLevel &L)
{
    grmp_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    double *diag = nullptr;
    GRMP_TRY(amg_alloc(m, &L.val32, L.nnz));
    GRMP_TRY(amg_alloc(m, &L.dinv, L.n));
    GRMP_TRY(amg_alloc(m, &L.lam, 1));
    GRMP_TRY(dalloc(ctx, &diag, L.n));
    hipLaunchKernelGGL(k_amg_diag, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.rp, L.ci, L.val, L.n, diag);
    hipLaunchKernelGGL(k_amg_dinv, dim3(nblk(L.n, 256)), dim3(256), 0, st, diag, L.fixed, L.n, L.dinv);
    hipLaunchKernelGGL(k_mg_val32, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.rp, L.val, L.fixed, L.val32, L.n);
    const hipError_t e = hipMemsetAsync(L.lam, 0, 8, st);
    hipLaunchKernelGGL(k_mg_gersh, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.rp, L.val, L.dinv, L.fixed, L.n,
                       (unsigned long long *)L.lam);
    dfree(ctx, diag);
    GRMP_HIP(ctx, e);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

static void amg_free(grmp_amg *m)
{
    if (!m) return;
    for (void *p : m->mem) dfree(m->ctx, p);
    m->mem.clear();
    if (m->h) (void)hipHostFree(m->h);
    if (m->ev) (void)hipEventDestroy(m->ev);
    delete m;
}

static int amg_build(grmp_amg *m)
{
    grmp_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    const grmp_matrix *A = m->A;
    AmgLevel &L0 = m->lv[0];
    L0.n = A->n;
    L0.nnz = A->nnz;
    L0.rp = A->rowptr;
    L0.ci = A->cols;
    L0.val = A->val64;
    L0.fixed = A->fixed;
    m->nlev = 1;
    m->nnz_total = L0.nnz;
    GRMP_TRY(amg_level_smoother(m, L0));
    for (;;) {
        AmgLevel &F = m->lv[m->nlev - 1];
        // coarsest once few enough free rows remain (the dense inverse replaces fixed rows by the identity)
        if (F.n <= MG_DENSE_MAX) break;
        GRMP_CHECK(ctx, m->nlev < AMG_MAX_LEVELS, GRMP_ERR_UNSUPPORTED,
                   "AMG: %lld rows left after %d levels", (long long)F.n, m->nlev);
        // first pairwise pass on the level, second on the intermediate Galerkin product
        int32_t *a1 = nullptr, *a2 = nullptr, *irp = nullptr, *ici = nullptr;
        double *ival = nullptr;
        int64_t n1 = 0, n2 = 0, innz = 0;
        int rc = GRMP_OK;
        do {
            if ((rc = dalloc(ctx, &a1, F.n))) break;
            if ((rc = amg_pairwise(ctx, F.rp, F.ci, F.val, F.fixed, F.n, a1, &n1))) break;
            if (n1 == 0) {
                rc = grmp_fail(ctx, GRMP_ERR_ARG, "AMG: no free rows");
                break;
            }
            if ((rc = amg_galerkin(ctx, F.rp, F.ci, F.val, F.n, F.nnz, a1, n1, &irp, &ici, &ival, &innz))) break;
            if ((rc = dalloc(ctx, &a2, n1))) break;
            if ((rc = amg_pairwise(ctx, irp, ici, ival, nullptr, n1, a2, &n2))) break;
        } while (0);
        dfree(ctx, irp);
        dfree(ctx, ici);
        dfree(ctx, ival);
        if (rc) {
            dfree(ctx, a1);
            dfree(ctx, a2);
            return rc;
        }
        if (n2 > F.n - F.n / 10) {  // coarsening stalled
            dfree(ctx, a1);
            dfree(ctx, a2);
            return grmp_fail(ctx, GRMP_ERR_UNSUPPORTED, "AMG: coarsening stalled at %lld rows", (long long)F.n);
        }
        rc = amg_alloc(m, &F.agg, F.n);
        if (!rc) hipLaunchKernelGGL(k_amg_compose, dim3(nblk(F.n, 256)), dim3(256), 0, st, a1, a2, F.n, F.agg);
        dfree(ctx, a1);
        dfree(ctx, a2);
        GRMP_TRY(rc);
        // the coarse level
        AmgLevel &C = m->lv[m->nlev];
        C.n = n2;
        int32_t *crp = nullptr, *cci = nullptr;
        double *cval = nullptr;
        GRMP_TRY(amg_galerkin(ctx, F.rp, F.ci, F.val, F.n, F.nnz, F.agg, n2, &crp, &cci, &cval, &C.nnz));
        m->mem.push_back(crp);
        m->mem.push_back(cci);
        m->mem.push_back(cval);
        C.rp = crp;
        C.ci = cci;
        C.val = cval;
        uint8_t *cfix = nullptr;
        GRMP_TRY(amg_alloc(m, &cfix, C.n));
        GRMP_HIP(ctx, hipMemsetAsync(cfix, 0, C.n, st));
        C.fixed = cfix;
        // members of each coarse row, in fine-row order, for the restriction
        {
            int32_t *key = nullptr, *idx = nullptr, *skey = nullptr, *cnt = nullptr;
            int rc2 = GRMP_OK;
            do {
                if ((rc2 = dalloc(ctx, &key, F.n)) || (rc2 = dalloc(ctx, &idx, F.n)) || (rc2 = dalloc(ctx, &skey, F.n)) ||
                    (rc2 = dalloc(ctx, &cnt, C.n)) || (rc2 = amg_alloc(m, &C.midx, F.n)) ||
                    (rc2 = amg_alloc(m, &C.mptr, C.n + 1)))
                    break;
                hipLaunchKernelGGL(k_amg_member_keys, dim3(nblk(F.n, 256)), dim3(256), 0, st, F.agg, F.n, (int32_t)C.n,
                                   key, idx);
                int bits = 1;
                while (bits < 31 && ((int64_t)1 << bits) <= C.n) ++bits;
                if ((rc2 = sort_pairs_i32_i32(ctx, key, skey, idx, C.midx, F.n, bits))) break;
                if (hipMemsetAsync(cnt, 0, sizeof(int32_t) * C.n, st) != hipSuccess) {
                    rc2 = grmp_fail(ctx, GRMP_ERR_HIP, "AMG members: memset");
                    break;
                }
                hipLaunchKernelGGL(k_amg_member_count, dim3(nblk(F.n, 256)), dim3(256), 0, st, F.agg, F.n, cnt);
                rc2 = exclusive_scan_i32(ctx, cnt, C.mptr, C.n);
            } while (0);
            dfree(ctx, key);
            dfree(ctx, idx);
            dfree(ctx, skey);
            dfree(ctx, cnt);
            GRMP_TRY(rc2);
        }
        ++m->nlev;
        m->nnz_total += C.nnz;
        GRMP_TRY(amg_level_smoother(m, C));
    }
    // vectors: level 0 aliases the CG's z and r; coarse levels own theirs
    const int64_t n0 = L0.n;
    m->nb0 = (int)nblk(n0, KB);
    GRMP_TRY(amg_alloc(m, &m->x, n0));
    GRMP_TRY(amg_alloc(m, &m->rhs, n0));
    GRMP_TRY(amg_alloc(m, &m->r, n0));
    GRMP_TRY(amg_alloc(m, &m->z, n0));
    GRMP_TRY(amg_alloc(m, &m->p, n0));
    GRMP_TRY(amg_alloc(m, &m->q, n0));
    GRMP_TRY(amg_alloc(m, &m->part, m->nb0));
    GRMP_TRY(amg_alloc(m, &m->part2, m->nb0));
    GRMP_TRY(amg_alloc(m, &m->sc, 1));
    for (int l = 0; l < m->nlev; ++l) {
        AmgLevel &L = m->lv[l];
        if (l == 0) {
            L.x = m->z;
            L.b = m->r;
        } else {
            GRMP_TRY(amg_alloc(m, &L.x, L.n));
            GRMP_TRY(amg_alloc(m, &L.b, L.n));
        }
        GRMP_TRY(amg_alloc(m, &L.r, L.n));
        GRMP_TRY(amg_alloc(m, &L.d, L.n));
    }
    const AmgLevel &C = m->lv[m->nlev - 1];
    GRMP_TRY(amg_alloc(m, &m->inv, C.n * C.n));
    hipLaunchKernelGGL(k_mg_dense_inv, dim3(1), dim3(256), 0, st, C.rp, C.ci, C.val, C.fixed, (int)C.n, m->inv);
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_HIP(ctx, hipStreamSynchronize(st));
    return GRMP_OK;
}

static void amg_rows(MgRows &a, const AmgLevel &L, const grmp_amg *m)
{
    memset(&a, 0, sizeof(a));
    a.rp = L.rp;
    a.ci = L.ci;
    a.val = L.val;
    a.val32 = L.val32;
    a.fixed = L.fixed;
    a.n = L.n;
    a.dinv = L.dinv;
    a.lam = L.lam;
    a.sc = m->sc;
}

// z (= x_0) = V(r (= b_0)); level-0 partials of r.z land in part
static void amg_vcycle(grmp_amg *m)
{
    hipStream_t st = m->ctx->stream;
    const int L1 = m->nlev - 1;
    for (int l = 0; l < L1; ++l) {
        const AmgLevel &L = m->lv[l], &C = m->lv[l + 1];
        MgRows a;
        amg_rows(a, L, m);
        a.b = L.b;
        a.o1 = L.x;
        hipLaunchKernelGGL(k_mg_pre, dim3(nblk(L.n, KB)), dim3(KB), 0, st, a);
        amg_rows(a, L, m);
        a.v = L.x;
        a.b = L.b;
        a.o1 = L.r;
        hipLaunchKernelGGL(k_mg_rows<OP_RESID>, dim3(nblk(L.n, KB)), dim3(KB), 0, st, a);
        hipLaunchKernelGGL(k_amg_restrict, dim3(nblk(C.n, 256)), dim3(256), 0, st, C.mptr, C.midx, L.r, C.b, C.n,
                           m->sc);
    }
    const AmgLevel &C = m->lv[L1];
    hipLaunchKernelGGL(k_mg_dense_apply, dim3(1), dim3(MG_DENSE_MAX), 0, st, m->inv, C.b, (int)C.n, C.x, m->sc);
    for (int l = L1 - 1; l >= 0; --l) {
        const AmgLevel &L = m->lv[l], &Cc = m->lv[l + 1];
        hipLaunchKernelGGL(k_amg_prolong, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.agg, Cc.x, AMG_SCALE, L.x, L.n,
                           m->sc);
        MgRows a;
        amg_rows(a, L, m);
        a.v = L.x;
        a.b = L.b;
        a.o1 = L.d;
        a.o2 = L.r;
        hipLaunchKernelGGL(k_mg_rows<OP_CHEB_POST1>, dim3(nblk(L.n, KB)), dim3(KB), 0, st, a);
        amg_rows(a, L, m);
        a.v = L.r;
        a.b = L.b;
        a.d = L.d;
        a.o1 = L.x;
        a.part = l == 0 ? m->part : nullptr;
        hipLaunchKernelGGL(k_mg_rows<OP_CHEB_POST2>, dim3(nblk(L.n, KB)), dim3(KB), 0, st, a);
    }
    if (L1 == 0) {  // one level: z = the dense solve, partials of r.z
        hipLaunchKernelGGL(k_mg_dot, dim3(m->nb0), dim3(KB), 0, st, m->r, m->z, C.n, m->part, m->sc);
    }
}

static void amg_iterate(grmp_amg *m, int iters)
{
    hipStream_t st = m->ctx->stream;
    const AmgLevel &L0 = m->lv[0];
    const int64_t n = L0.n;
    for (int k = 0; k < iters; ++k) {
        MgRows a;
        amg_rows(a, L0, m);
        a.v = m->p;
        a.o1 = m->q;
        a.part = m->part;
        hipLaunchKernelGGL(k_mg_rows<OP_PCG_Q>, dim3(nblk(n, KB)), dim3(KB), 0, st, a);
        hipLaunchKernelGGL(k_mg_red<RED_ALPHA>, dim3(1), dim3(KB), 0, st, m->part, m->nb0, nullptr, m->sc);
        hipLaunchKernelGGL(k_mg_axpy, dim3(m->nb0), dim3(KB), 0, st, m->x, m->r, m->p, m->q, n, m->part, m->sc);
        hipLaunchKernelGGL(k_mg_red<RED_CONV>, dim3(1), dim3(KB), 0, st, m->part, m->nb0, nullptr, m->sc);
        amg_vcycle(m);
        hipLaunchKernelGGL(k_mg_red<RED_BETA>, dim3(1), dim3(KB), 0, st, m->part, m->nb0, nullptr, m->sc);
        hipLaunchKernelGGL(k_mg_update_p, dim3(nblk(n, 256)), dim3(256), 0, st, m->p, m->z, n, m->sc);
    }
}

// the true residual b - A x (FP64 values): accept, or restart CG from it
static void amg_verify(grmp_amg *m)
{
    hipStream_t st = m->ctx->stream;
    const AmgLevel &L0 = m->lv[0];
    const int64_t n = L0.n;
    hipLaunchKernelGGL(k_mg_verify_begin, dim3(1), dim3(1), 0, st, m->sc);
    MgRows a;
    amg_rows(a, L0, m);
    a.v = m->x;
    a.b = m->rhs;
    a.o1 = m->r;
    hipLaunchKernelGGL(k_mg_rows<OP_TRUE_RESID>, dim3(nblk(n, KB)), dim3(KB), 0, st, a);
    hipLaunchKernelGGL(k_mg_dot, dim3(m->nb0), dim3(KB), 0, st, m->r, m->r, n, m->part, m->sc);
    hipLaunchKernelGGL(k_mg_red<RED_VERIFY>, dim3(1), dim3(KB), 0, st, m->part, m->nb0, nullptr, m->sc);
    amg_vcycle(m);
    hipLaunchKernelGGL(k_mg_red<RED_RESTART>, dim3(1), dim3(KB), 0, st, m->part, m->nb0, nullptr, m->sc);
    hipLaunchKernelGGL(k_mg_update_p, dim3(nblk(n, 256)), dim3(256), 0, st, m->p, m->z, n, m->sc);
}

static int amg_status(grmp_amg *m)
{
    hipStream_t st = m->ctx->stream;
    GRMP_HIP(m->ctx, hipMemcpyAsync(&m->h->relres, &m->sc->relres, 8, hipMemcpyDeviceToHost, st));
    GRMP_HIP(m->ctx, hipMemcpyAsync(&m->h->it, &m->sc->it, 12, hipMemcpyDeviceToHost, st));
    GRMP_HIP(m->ctx, hipEventRecord(m->ev, st));
    GRMP_HIP(m->ctx, hipEventSynchronize(m->ev));
    return GRMP_OK;
}

extern "C" int grmp_amg_destroy(grmp_amg *m)
{
    if (!m) return GRMP_OK;
    grmp_ctx *ctx = m->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipStreamSynchronize(ctx->stream);
    amg_free(m);
    return GRMP_OK;
}

// z = V v: one V-cycle of the hierarchy from a zero start (fixed rows 0), as
// the right preconditioner of BiCGStab on a matrix of the hierarchy's rows,
// e.g. the Newton matrix it was built from
int amg_precond_apply(grmp_amg *m, const double *v, double *z, hipStream_t st)
{
    GRMP_CHECK(m->ctx, st == m->ctx->stream, GRMP_ERR_ARG,
               "algebraic multigrid preconditioner: not the hierarchy's stream");
    const AmgLevel &L0 = m->lv[0];
    hipLaunchKernelGGL(k_mg_pc_in, dim3(nblk(L0.n, 256)), dim3(256), 0, st, v, L0.fixed, m->r, m->sc, L0.n);
    amg_vcycle(m);
    GRMP_HIP(m->ctx, hipMemcpyAsync(z, m->z, sizeof(double) * L0.n, hipMemcpyDeviceToDevice, st));
    GRMP_HIP(m->ctx, hipGetLastError());
    return GRMP_OK;
}

int64_t amg_level0_rows(const grmp_amg *m) { return m->lv[0].n; }
grmp_ctx *amg_context(const grmp_amg *m) { return m->ctx; }

extern "C" int grmp_amg_create(grmp_matrix *A, grmp_amg **out)
{
    if (!A || !out) return GRMP_ERR_ARG;
    *out = nullptr;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, A->have_pattern && A->val64 && A->fixed && A->n >= 1, GRMP_ERR_ARG,
               "grmp_amg_create: matrix not assembled in FP64");
    std::lock_guard<std::mutex> g(ctx->mu);
    grmp_amg *m = new grmp_amg();
    m->ctx = ctx;
    m->A = A;
    int rc = GRMP_OK;
    if (hipHostMalloc((void **)&m->h, sizeof(MgStatus), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev, hipEventDisableTiming) != hipSuccess) {
        m->h = nullptr;
        rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_amg_create: status buffer");
    }
    if (!rc) rc = amg_build(m);
    if (rc) {
        (void)hipStreamSynchronize(ctx->stream);
        amg_free(m);
        return rc;
    }
    *out = m;
    return GRMP_OK;
}

extern "C" int grmp_amg_info(const grmp_amg *m, int *nlev, int64_t *coarse_rows, double *complexity)
{
    if (!m) return GRMP_ERR_ARG;
    if (nlev) *nlev = m->nlev;
    if (coarse_rows) *coarse_rows = m->lv[m->nlev - 1].n;
    if (complexity) *complexity = (double)m->nnz_total / (double)std::max<int64_t>(m->lv[0].nnz, 1);
    return GRMP_OK;
}

extern "C" int grmp_amg_apply(grmp_amg *m, const grmp_vector *b, grmp_vector *x, double rtol, int64_t maxit,
                              grmp_solve_info *info)
{
    if (!m || !b || !x) return GRMP_ERR_ARG;
    grmp_ctx *ctx = m->ctx;
    const int64_t n = m->lv[0].n;
    GRMP_CHECK(ctx, b->ctx == ctx && x->ctx == ctx && b->n == n && x->n == n, GRMP_ERR_ARG,
               "grmp_amg_apply: b and x need length %lld", (long long)n);
    GRMP_CHECK(ctx, rtol > 0.0 && rtol < 1.0 && maxit >= 1 && maxit <= (1 << 30), GRMP_ERR_ARG,
               "grmp_amg_apply: rtol in (0, 1) and maxit >= 1 needed");
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = ctx->stream;
    const double t0 = wall_seconds();
    MgStatus *h = m->h;
    memset(h, 0, sizeof(MgStatus));
    int rc = GRMP_OK;
    do {
        // x = b_d / A_dd on fixed rows, 0 elsewhere; r = b - A x on free rows; z = V(r); p = z
        MgScal s0;
        memset(&s0, 0, sizeof(s0));
        s0.rtol = rtol;
        s0.maxit = (int)maxit;
        if (hipMemcpyAsync(m->sc, &s0, sizeof(s0), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(m->rhs, b->d, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemsetAsync(m->p, 0, sizeof(double) * n, st) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_amg_apply: start");
            break;
        }
        if ((rc = init_x_launch(m->A, m->A->val64, m->rhs, m->x, st))) break;
        MgRows a;
        amg_rows(a, m->lv[0], m);
        a.v = m->x;
        a.b = m->rhs;
        a.o1 = m->r;
        hipLaunchKernelGGL(k_mg_rows<OP_TRUE_RESID>, dim3(nblk(n, KB)), dim3(KB), 0, st, a);
        amg_vcycle(m);
        hipLaunchKernelGGL(k_mg_dot, dim3(m->nb0), dim3(KB), 0, st, m->r, m->r, n, m->part2, m->sc);
        hipLaunchKernelGGL(k_mg_red<RED_INIT>, dim3(1), dim3(KB), 0, st, m->part, m->nb0, m->part2, m->sc);
        hipLaunchKernelGGL(k_mg_update_p, dim3(nblk(n, 256)), dim3(256), 0, st, m->p, m->z, n, m->sc);
        amg_iterate(m, 4);
        if ((rc = amg_status(m))) break;
        while (h->done != 1) {
            if (h->done == 2) amg_verify(m);
            else amg_iterate(m, 4);
            if ((rc = amg_status(m))) break;
        }
        if (rc) break;
        if (hipMemcpyAsync(x->d, m->x, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_amg_apply: solution");
    } while (0);
    grmp_solve_info inf;
    memset(&inf, 0, sizeof(inf));
    inf.status = h->status;
    inf.iterations = h->it;
    inf.relres = h->relres;
    inf.outer = 1;
    inf.seconds = wall_seconds() - t0;
    if (info) *info = inf;
    if (rc) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    if (inf.status == GRMP_ERR_NOCONV)
        return grmp_fail(ctx, GRMP_ERR_NOCONV, "grmp_amg_apply: no convergence, relres %.3e after %lld iterations",
                         inf.relres, (long long)inf.iterations);
    return inf.status;
}
