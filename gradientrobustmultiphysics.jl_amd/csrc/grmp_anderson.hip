// Anderson acceleration of the fixpoint loop: AndersonAccelerationManager
// and update_anderson! (src/solvers.jl:688-808), called by
// solve_fixpoint_full! after each linear solve (src/solvers.jl:923-929).
//
// The iterates live on the device in two rings of depth + 1 vectors: the
// fixpoint outputs ("tilde" iterates) and the accelerated iterates.  The
// reference shifts its arrays by copies; here a shift moves the ring's start
// index, which leaves every logical slot with the same contents.  The
// accelerated ring shifts only once the depth counter is positive, as in the
// reference, so the second fixpoint output is paired with the initial iterate.
//
// Per update with depth counter c > 0:
//   d_j = (tilde_j - last_j) on the accelerated entries,  j = 1 .. c+1
//   G_jk = d_j . (B d_k)   (B = identity for "l2", else the metric matrix)
//   [G 1; 1' 0] alpha = [0; 1]   in the reference's (depth + 2) system
//   target = sum_a damping alpha_a tilde_a + (1 - damping) alpha_a last_a
// The Gram entries are block-partial sums folded in a fixed order (the same
// bits on every run); the small system is solved on the host by Gaussian
// elimination with partial pivoting, as Julia's dense `\` does.  This file is
// compiled without contraction so that the combination rounds as the
// reference's loop does for the same alpha.
#include "grmp_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

#define AA_MAX_DEPTH 10
#define AA_BLOCK 256
#define AA_MAX_BLOCKS 512

struct AaVecs {
    const double *t[AA_MAX_DEPTH + 1];  // tilde iterate of pair j (0-based)
    const double *l[AA_MAX_DEPTH + 1];  // accelerated iterate of pair j
    double ct[AA_MAX_DEPTH + 1];        // damping * alpha_j
    double cl[AA_MAX_DEPTH + 1];        // (1 - damping) * alpha_j
};

struct grmp_anderson {
    grmp_ctx *ctx = nullptr;
    int64_t n = 0;
    int m = 0;                    // anderson_iterations
    double damping = 1.0;
    const grmp_matrix *metric = nullptr;  // nullptr: the l2 metric
    uint8_t *inc = nullptr;       // [n] 1 on the accelerated entries
    double *ring = nullptr;       // [2 (m+1) n]: tilde slots, then accelerated slots
    double *dif = nullptr;        // [(m+1) n] differences d_j
    double *bd = nullptr;         // [(m+1) n] B d_j (metric only)
    double *part = nullptr;       // [npairs * nb] block partials
    double *gram = nullptr;       // [npairs]
    double *hgram = nullptr;      // pinned [npairs]
    int tstart = 0, lstart = 0;   // ring starts (logical slot s lives in physical (start + s) % (m+1))
    int cdepth = -1;
    int nb = 1;
    std::vector<double> M, rhs, alpha;  // the reference's (m+2) x (m+2) system, row-major
};

static double *aa_tilde(grmp_anderson *a, int s) { return a->ring + (size_t)((a->tstart + s) % (a->m + 1)) * a->n; }
static double *aa_last(grmp_anderson *a, int s)
{
    return a->ring + (size_t)(a->m + 1 + (a->lstart + s) % (a->m + 1)) * a->n;
}

__global__ void k_aa_diff(AaVecs v, int np, const uint8_t *__restrict__ inc, int64_t n, double *__restrict__ dif)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool on = inc[i] != 0;
    for (int j = 0; j < np; ++j) dif[(size_t)j * n + i] = on ? v.t[j][i] - v.l[j][i] : 0.0;
}

// partial sums of d_j . w_k over one contiguous chunk per block, pair p = blockIdx.y
__global__ __launch_bounds__(AA_BLOCK) void k_aa_gram_part(const double *__restrict__ dif, const double *__restrict__ w,
                                                           int np, int64_t n, int64_t chunk, double *__restrict__ part)
{
    __shared__ double red[AA_BLOCK];
    int p = blockIdx.y, j = 0;
    while (p >= np - j) {  // pair p -> (j, k), j <= k, rows of the upper triangle in order
        p -= np - j;
        ++j;
    }
    const int k = j + p;
    const double *a = dif + (size_t)j * n, *b = w + (size_t)k * n;
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += AA_BLOCK) s += a[i] * b[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = AA_BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(AA_BLOCK) void k_aa_gram_fold(const double *__restrict__ part, int nb,
                                                           double *__restrict__ gram)
{
    __shared__ double red[AA_BLOCK];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += AA_BLOCK) s += part[(size_t)blockIdx.x * nb + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = AA_BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) gram[blockIdx.x] = red[0];
}

// target_i = sum_a ct_a tilde_a,i + cl_a last_a,i on the accelerated entries,
// summed from zero in the reference's order (src/solvers.jl:785-794)
__global__ void k_aa_combine(AaVecs v, int np, const uint8_t *__restrict__ inc, int64_t n, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !inc[i]) return;
    double s = 0.0;
    for (int a = 0; a < np; ++a) {
        s = s + v.ct[a] * v.t[a][i];
        s = s + v.cl[a] * v.l[a][i];
    }
    x[i] = s;
}

static void aa_free(grmp_anderson *a)
{
    if (!a) return;
    grmp_ctx *ctx = a->ctx;
    dfree(ctx, a->inc);
    dfree(ctx, a->ring);
    dfree(ctx, a->dif);
    dfree(ctx, a->bd);
    dfree(ctx, a->part);
    dfree(ctx, a->gram);
    if (a->hgram) (void)hipHostFree(a->hgram);
    (void)hipStreamSynchronize(ctx->stream);
    delete a;
}

static int aa_init(grmp_anderson *a, int64_t nranges, const int64_t *ranges, const grmp_vector *initial)
{
    grmp_ctx *ctx = a->ctx;
    hipStream_t st = ctx->stream;
    const int64_t n = a->n;
    const int np = a->m + 1;
    const int npairs = np * (np + 1) / 2;
    int64_t chunk = (n + AA_MAX_BLOCKS - 1) / AA_MAX_BLOCKS;
    chunk = (chunk + AA_BLOCK - 1) / AA_BLOCK * AA_BLOCK;
    a->nb = (int)((n + chunk - 1) / chunk);
    GRMP_TRY(dalloc(ctx, &a->inc, n));
    GRMP_TRY(dalloc(ctx, &a->ring, (size_t)2 * np * n));
    GRMP_TRY(dalloc(ctx, &a->dif, (size_t)np * n));
    if (a->metric) GRMP_TRY(dalloc(ctx, &a->bd, (size_t)np * n));
    GRMP_TRY(dalloc(ctx, &a->part, (size_t)npairs * a->nb));
    GRMP_TRY(dalloc(ctx, &a->gram, npairs));
    GRMP_HIP(ctx, hipHostMalloc((void **)&a->hgram, sizeof(double) * npairs, hipHostMallocDefault));
    std::vector<uint8_t> h(n, 0);
    for (int64_t r = 0; r < nranges; ++r)
        for (int64_t i = ranges[2 * r]; i < ranges[2 * r + 1]; ++i) h[i] = 1;
    GRMP_HIP(ctx, hipMemcpyAsync(a->inc, h.data(), n, hipMemcpyHostToDevice, st));
    // every slot of both rings starts as the initial iterate (src/solvers.jl:740-744)
    for (int s = 0; s < 2 * np; ++s)
        GRMP_HIP(ctx, hipMemcpyAsync(a->ring + (size_t)s * n, initial->d, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                     st));
    const int K = a->m + 2;
    a->M.assign((size_t)K * K, 0.0);
    a->rhs.assign(K, 0.0);
    a->alpha.assign(K, 0.0);
    a->rhs[K - 1] = 1.0;  // the coefficients sum to one
    for (int j = 0; j < np; ++j) a->M[(size_t)j * K + j] = 1.0;
    GRMP_HIP(ctx, hipStreamSynchronize(st));
    return GRMP_OK;
}

extern "C" int grmp_anderson_create(grmp_ctx *ctx, int64_t n, int depth, double damping, const grmp_matrix *metric,
                                    int64_t nranges, const int64_t *ranges, const grmp_vector *initial,
                                    grmp_anderson **out)
{
    if (!ctx || !out) return GRMP_ERR_ARG;
    *out = nullptr;
    GRMP_CHECK(ctx, depth > 0 && depth <= AA_MAX_DEPTH, GRMP_ERR_ARG,
               "grmp_anderson_create: anderson_iterations %d not in 1..%d", depth, AA_MAX_DEPTH);
    GRMP_CHECK(ctx, damping > 0.0 && damping <= 1.0, GRMP_ERR_ARG,
               "grmp_anderson_create: damping %g not in (0, 1]", damping);
    GRMP_CHECK(ctx, n > 0 && initial && initial->ctx == ctx && initial->n == n, GRMP_ERR_ARG,
               "grmp_anderson_create: initial iterate missing or not of length %lld", (long long)n);
    GRMP_CHECK(ctx, !metric || (metric->ctx == ctx && metric->n == n && metric->have_pattern && metric->val64),
               GRMP_ERR_ARG, "grmp_anderson_create: metric matrix not assembled in FP64 with %lld rows", (long long)n);
    GRMP_CHECK(ctx, nranges > 0 && ranges, GRMP_ERR_ARG, "grmp_anderson_create: no entry ranges");
    for (int64_t r = 0; r < nranges; ++r)
        GRMP_CHECK(ctx, 0 <= ranges[2 * r] && ranges[2 * r] <= ranges[2 * r + 1] && ranges[2 * r + 1] <= n,
                   GRMP_ERR_ARG, "grmp_anderson_create: range %lld [%lld, %lld) outside [0, %lld)", (long long)r,
                   (long long)ranges[2 * r], (long long)ranges[2 * r + 1], (long long)n);
    std::lock_guard<std::mutex> g(ctx->mu);
    grmp_anderson *a = new grmp_anderson();
    a->ctx = ctx;
    a->n = n;
    a->m = depth;
    a->damping = damping;
    a->metric = metric;
    const int rc = aa_init(a, nranges, ranges, initial);
    if (rc) {
        aa_free(a);
        return rc;
    }
    *out = a;
    return GRMP_OK;
}

// Gaussian elimination with partial pivoting on a copy of the K x K system
static bool aa_dense_solve(std::vector<double> A, std::vector<double> b, int K, std::vector<double> &x)
{
    for (int c = 0; c < K; ++c) {
        int p = c;
        for (int r = c + 1; r < K; ++r)
            if (std::fabs(A[(size_t)r * K + c]) > std::fabs(A[(size_t)p * K + c])) p = r;
        if (A[(size_t)p * K + c] == 0.0) return false;
        if (p != c) {
            for (int k = 0; k < K; ++k) std::swap(A[(size_t)p * K + k], A[(size_t)c * K + k]);
            std::swap(b[p], b[c]);
        }
        for (int r = c + 1; r < K; ++r) {
            const double f = A[(size_t)r * K + c] / A[(size_t)c * K + c];
            if (f == 0.0) continue;
            for (int k = c; k < K; ++k) A[(size_t)r * K + k] -= f * A[(size_t)c * K + k];
            b[r] -= f * b[c];
        }
    }
    x.assign(K, 0.0);
    for (int r = K - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < K; ++k) s -= A[(size_t)r * K + k] * x[k];
        x[r] = s / A[(size_t)r * K + r];
    }
    return true;
}

static int aa_update(grmp_anderson *a, grmp_vector *target, double *alpha_out, double *local_residual)
{
    grmp_ctx *ctx = a->ctx;
    hipStream_t st = ctx->stream;
    const int64_t n = a->n;
    const int m = a->m, K = m + 2;
    const int cd = a->cdepth < m ? a->cdepth : m;
    const unsigned gb = (unsigned)((n + 255) / 256);
    if (local_residual) *local_residual = 0.0;
    if (alpha_out)
        for (int j = 0; j <= m; ++j) alpha_out[j] = 0.0;
    // shift the tilde ring and keep the fixpoint output as its newest slot
    a->tstart = (a->tstart + 1) % (m + 1);
    GRMP_HIP(ctx, hipMemcpyAsync(aa_tilde(a, m), target->d, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if (cd > 0) {
        const int np = cd + 1, npairs = np * (np + 1) / 2;
        AaVecs v;
        memset(&v, 0, sizeof(v));
        for (int j = 0; j < np; ++j) {  // pair j + 1 of the reference: slots m+1-(j+1) (0-based m - j)
            v.t[j] = aa_tilde(a, m - j);
            v.l[j] = aa_last(a, m - j);
        }
        hipLaunchKernelGGL(k_aa_diff, dim3(gb), dim3(256), 0, st, v, np, a->inc, n, a->dif);
        const double *w = a->dif;
        if (a->metric) {
            for (int j = 0; j < np; ++j)
                GRMP_TRY(spmv_launch(ctx, a->metric, GRMP_VALUES_FP64, a->dif + (size_t)j * n, a->bd + (size_t)j * n,
                                     st));
            w = a->bd;
        }
        const int64_t chunk = (n + a->nb - 1) / a->nb;
        const int nb = (int)((n + chunk - 1) / chunk);
        hipLaunchKernelGGL(k_aa_gram_part, dim3(nb, npairs), dim3(AA_BLOCK), 0, st, a->dif, w, np, n, chunk, a->part);
        hipLaunchKernelGGL(k_aa_gram_fold, dim3(npairs), dim3(AA_BLOCK), 0, st, a->part, nb, a->gram);
        GRMP_HIP(ctx, hipGetLastError());
        GRMP_HIP(ctx, hipMemcpyAsync(a->hgram, a->gram, sizeof(double) * npairs, hipMemcpyDeviceToHost, st));
        GRMP_HIP(ctx, hipStreamSynchronize(st));
        // fill the system up to the current depth (src/solvers.jl:764-772)
        for (int j = 0, p = 0; j < np; ++j)
            for (int k = j; k < np; ++k, ++p) {
                const double gjk = a->hgram[p];
                GRMP_CHECK(ctx, std::isfinite(gjk), GRMP_ERR_NONFINITE,
                           "grmp_anderson_update: non-finite Gram entry (%d, %d)", j + 1, k + 1);
                a->M[(size_t)j * K + k] = gjk;
                a->M[(size_t)k * K + j] = gjk;
                a->M[(size_t)j * K + (K - 1)] = 1.0;
                a->M[(size_t)(K - 1) * K + j] = 1.0;
            }
        GRMP_CHECK(ctx, aa_dense_solve(a->M, a->rhs, K, a->alpha), GRMP_ERR_NONFINITE,
                   "grmp_anderson_update: the local system is singular");
        double res = 0.0;
        for (int r = 0; r < K; ++r) {
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += a->M[(size_t)r * K + k] * a->alpha[k];
            res += (s - a->rhs[r]) * (s - a->rhs[r]);
        }
        if (local_residual) *local_residual = std::sqrt(res);
        for (int j = 0; j < np; ++j) {
            GRMP_CHECK(ctx, std::isfinite(a->alpha[j]), GRMP_ERR_NONFINITE,
                       "grmp_anderson_update: non-finite coefficient %d", j + 1);
            v.ct[j] = a->damping * a->alpha[j];
            v.cl[j] = (1.0 - a->damping) * a->alpha[j];
            if (alpha_out) alpha_out[j] = a->alpha[j];
        }
        hipLaunchKernelGGL(k_aa_combine, dim3(gb), dim3(256), 0, st, v, np, a->inc, n, target->d);
        GRMP_HIP(ctx, hipGetLastError());
        a->lstart = (a->lstart + 1) % (m + 1);  // shift the accelerated ring
    }
    // keep the new iterate (src/solvers.jl:802-803)
    GRMP_HIP(ctx, hipMemcpyAsync(aa_last(a, m), target->d, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    a->cdepth += 1;
    return GRMP_OK;
}

extern "C" int grmp_anderson_update(grmp_anderson *a, grmp_vector *target, double *alpha_out, double *local_residual)
{
    if (!a || !target) return GRMP_ERR_ARG;
    grmp_ctx *ctx = a->ctx;
    GRMP_CHECK(ctx, target->ctx == ctx && target->n == a->n, GRMP_ERR_ARG,
               "grmp_anderson_update: target not of length %lld in the manager's context", (long long)a->n);
    std::lock_guard<std::mutex> g(ctx->mu);
    return aa_update(a, target, alpha_out, local_residual);
}

extern "C" int grmp_anderson_depth(const grmp_anderson *a)
{
    return a ? a->cdepth : -2;
}

extern "C" int grmp_anderson_destroy(grmp_anderson *a)
{
    if (!a) return GRMP_OK;
    std::lock_guard<std::mutex> g(a->ctx->mu);
    aa_free(a);
    return GRMP_OK;
}
