// Context, device vectors, meshes and FE spaces of libgrmp_hip.so.
//
// The mesh and dofmap objects are the device-side images of ExtendableGrids'
// xgrid[Coordinates]/xgrid[CellNodes] and of FES[CellDofs]
// (src/finiteelements.jl:72-112, src/dofmaps.jl:201-363).  Everything is
// stored structure-of-arrays so that a thread per cell or per dof reads its
// entries with coalesced loads.
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstring>

#include "grmp_internal.h"

double wall_seconds()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int grmp_fail(grmp_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

int ensure_scratch(grmp_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return GRMP_OK;
    if (ctx->scratch) GRMP_HIP(ctx, hipFreeAsync(ctx->scratch, ctx->stream));
    ctx->scratch = nullptr;
    size_t nb = bytes + bytes / 2 + 4096;
    GRMP_HIP(ctx, hipMallocAsync(&ctx->scratch, nb, ctx->stream));
    ctx->scratch_bytes = nb;
    return GRMP_OK;
}

extern "C" int grmp_version(void) { return 100; }

// ------------------------------------------------------------ tuning table
// The autotuner writes one JSON object per entry, with integer fields
// "kind", "nq", "bucket", "variant", "rows", "remap" (other fields, such as
// names and timings, are ignored).  The reader takes every {...} object
// after "entries" that holds the three key fields.
static bool json_int(const std::string &obj, const char *key, long &out)
{
    const std::string k = std::string("\"") + key + "\"";
    size_t p = obj.find(k);
    if (p == std::string::npos) return false;
    p = obj.find(':', p + k.size());
    if (p == std::string::npos) return false;
    char *end = nullptr;
    out = std::strtol(obj.c_str() + p + 1, &end, 10);
    return end != obj.c_str() + p + 1;
}

static int tune_parse(grmp_ctx *ctx, const std::string &text, std::vector<TuneEntry> &out)
{
    size_t p = text.find("\"entries\"");
    GRMP_CHECK(ctx, p != std::string::npos, GRMP_ERR_ARG, "tuning table: no \"entries\" list");
    while ((p = text.find('{', p)) != std::string::npos) {
        const size_t q = text.find('}', p);
        if (q == std::string::npos) break;
        const std::string obj = text.substr(p, q - p + 1);
        long kind, nq, bucket, v = 0, rows = 0, remap = -1;
        if (json_int(obj, "kind", kind) && json_int(obj, "nq", nq) && json_int(obj, "bucket", bucket)) {
            json_int(obj, "variant", v);
            json_int(obj, "rows", rows);
            json_int(obj, "remap", remap);
            GRMP_CHECK(ctx, v >= 0 && v <= 3 && rows >= 0 && rows <= 1024 && remap >= -1 && remap <= 1, GRMP_ERR_ARG,
                       "tuning table: entry out of range in %s", obj.c_str());
            TuneEntry e;
            e.kind = (int)kind;
            e.nq = (int)nq;
            e.bucket = (int)bucket;
            e.t.variant = (int)v;
            e.t.rows = (int)rows;
            e.t.remap = (int)remap;
            out.push_back(e);
        }
        p = q + 1;
    }
    return GRMP_OK;
}

// the table shipped beside the library (tuning_gfx950.json), or $GRMP_TUNING
static std::string tune_default_path()
{
    if (const char *env = std::getenv("GRMP_TUNING")) return env;
    Dl_info info;
    if (dladdr((void *)&grmp_version, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t s = so.rfind('/');
        return (s == std::string::npos ? std::string(".") : so.substr(0, s)) + "/tuning_gfx950.json";
    }
    return "tuning_gfx950.json";
}

extern "C" int grmp_tune_load(grmp_ctx *ctx, const char *path)
{
    GRMP_CHECK(ctx, ctx, GRMP_ERR_ARG, "grmp_tune_load: null context");
    const std::string fn = path ? std::string(path) : tune_default_path();
    FILE *f = std::fopen(fn.c_str(), "rb");
    GRMP_CHECK(ctx, f, GRMP_ERR_ARG, "grmp_tune_load: cannot open %s", fn.c_str());
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    std::vector<TuneEntry> t;
    GRMP_TRY(tune_parse(ctx, text, t));
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->tune = t;
    return GRMP_OK;
}

extern "C" int grmp_tune_count(const grmp_ctx *ctx) { return ctx ? (int)ctx->tune.size() : -1; }

extern "C" int grmp_ctx_create(int device, grmp_ctx **out)
{
    if (!out) return GRMP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GRMP_ERR_HIP;
    if (device < 0 || device >= ndev) return GRMP_ERR_ARG;
    grmp_ctx *c = new grmp_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GRMP_ERR_HIP;
    }
    if (hipMalloc(&c->dflag, 64) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return GRMP_ERR_OOM;
    }
    // the shipped tuning table, when present (a missing file keeps the built-in launch shapes)
    const std::string fn = tune_default_path();
    if (FILE *f = std::fopen(fn.c_str(), "rb")) {
        std::fclose(f);
        if (grmp_tune_load(c, fn.c_str()) != GRMP_OK) c->tune.clear();
    }
    *out = c;
    return GRMP_OK;
}

extern "C" int grmp_ctx_destroy(grmp_ctx *ctx)
{
    if (!ctx) return GRMP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->dflag) (void)hipFree(ctx->dflag);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GRMP_OK;
}

extern "C" const char *grmp_last_error(const grmp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int grmp_ctx_synchronize(grmp_ctx *ctx)
{
    if (!ctx) return GRMP_ERR_ARG;
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GRMP_OK;
}

extern "C" int grmp_ctx_set_max_nnz(grmp_ctx *ctx, int64_t max_nnz)
{
    if (!ctx) return GRMP_ERR_ARG;
    GRMP_CHECK(ctx, max_nnz >= 1, GRMP_ERR_ARG, "grmp_ctx_set_max_nnz: the limit must be positive");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->max_nnz = max_nnz < 2147483647 ? max_nnz : 2147483647;
    return GRMP_OK;
}

int exclusive_scan_i32(grmp_ctx *ctx, const int32_t *in, int32_t *out, int64_t n)
{
    // out[0] = 0, out[i+1] = sum in[0..i]
    size_t tmp = 0;
    GRMP_HIP(ctx, rocprim::inclusive_scan(nullptr, tmp, in, out + 1, (size_t)n, rocprim::plus<int32_t>(),
                                          ctx->stream));
    GRMP_TRY(ensure_scratch(ctx, tmp));
    GRMP_HIP(ctx, hipMemsetAsync(out, 0, sizeof(int32_t), ctx->stream));
    GRMP_HIP(ctx, rocprim::inclusive_scan(ctx->scratch, tmp, in, out + 1, (size_t)n, rocprim::plus<int32_t>(),
                                          ctx->stream));
    return GRMP_OK;
}

int sum_i32_as_i64(grmp_ctx *ctx, const int32_t *in, int64_t n, int64_t *total)
{
    int64_t *d = nullptr;
    GRMP_TRY(dalloc(ctx, &d, 1));
    size_t tmp = 0;
    hipError_t e = rocprim::reduce(nullptr, tmp, in, d, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream);
    int rc = e == hipSuccess ? ensure_scratch(ctx, tmp) : GRMP_ERR_HIP;
    if (!rc) e = rocprim::reduce(ctx->scratch, tmp, in, d, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream);
    if (!rc && e == hipSuccess) e = hipMemcpyAsync(total, d, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, d);
    if (!rc && e != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "64-bit sum: %s", hipGetErrorString(e));
    return rc;
}

// ------------------------------------------ sorts and run sums (rocPRIM)
// Stable LSD radix sorts on the low `bits` key bits and the sums of runs of
// equal keys, used by the algebraic multigrid setup.  They run on the
// context's stream with its shared temporary storage.
int sort_pairs_u64_f64(grmp_ctx *ctx, const uint64_t *kin, uint64_t *kout, const double *vin, double *vout,
                       int64_t n, int bits)
{
    if (n <= 0) return GRMP_OK;
    size_t tmp = 0;
    GRMP_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits,
                                            ctx->stream));
    GRMP_TRY(ensure_scratch(ctx, tmp));
    GRMP_HIP(ctx, rocprim::radix_sort_pairs(ctx->scratch, tmp, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits,
                                            ctx->stream));
    return GRMP_OK;
}

int sort_pairs_i32_i32(grmp_ctx *ctx, const int32_t *kin, int32_t *kout, const int32_t *vin, int32_t *vout,
                       int64_t n, int bits)
{
    if (n <= 0) return GRMP_OK;
    size_t tmp = 0;
    GRMP_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits,
                                            ctx->stream));
    GRMP_TRY(ensure_scratch(ctx, tmp));
    GRMP_HIP(ctx, rocprim::radix_sort_pairs(ctx->scratch, tmp, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits,
                                            ctx->stream));
    return GRMP_OK;
}

int reduce_runs_u64_f64(grmp_ctx *ctx, const uint64_t *keys, const double *vals, int64_t n, uint64_t *ukeys,
                        double *sums, int64_t *count)
{
    *count = 0;
    if (n <= 0) return GRMP_OK;
    int64_t *dcount = nullptr;
    GRMP_TRY(dalloc(ctx, &dcount, 1));
    size_t tmp = 0;
    hipError_t e = rocprim::reduce_by_key(nullptr, tmp, keys, vals, (size_t)n, ukeys, sums, dcount,
                                          rocprim::plus<double>(), rocprim::equal_to<uint64_t>(), ctx->stream);
    int rc = e == hipSuccess ? ensure_scratch(ctx, tmp) : GRMP_ERR_HIP;
    if (!rc)
        e = rocprim::reduce_by_key(ctx->scratch, tmp, keys, vals, (size_t)n, ukeys, sums, dcount,
                                   rocprim::plus<double>(), rocprim::equal_to<uint64_t>(), ctx->stream);
    if (!rc && e == hipSuccess) e = hipMemcpyAsync(count, dcount, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, dcount);
    if (!rc && e != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "run sums: %s", hipGetErrorString(e));
    return rc;
}

// ------------------------------------------------------------ Philox RNG
// Philox-4x32-10 (Salmon et al., SC'11): counter = (index, 0, 0, 0),
// key = (seed low, seed high).
__device__ __host__ inline void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double u01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

__global__ void k_lognormal(double *v, int64_t n, uint32_t k0, uint32_t k1, double mu, double sigma)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u};
    philox4x32(c, k0, k1);
    double u1 = u01(c[0]), u2 = u01(c[1]);
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    v[i] = exp(mu + sigma * z);
}

__global__ void k_fill(double *v, int64_t n, double a)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------- vectors
extern "C" int grmp_vector_create(grmp_ctx *ctx, int64_t n, grmp_vector **out)
{
    GRMP_CHECK(ctx, ctx && out && n >= 0, GRMP_ERR_ARG, "grmp_vector_create: bad arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    grmp_vector *v = new grmp_vector{ctx, n, nullptr};
    int rc = dalloc(ctx, &v->d, (size_t)n);
    if (rc) { delete v; return rc; }
    GRMP_HIP(ctx, hipMemsetAsync(v->d, 0, sizeof(double) * (n ? n : 1), ctx->stream));
    *out = v;
    return GRMP_OK;
}

extern "C" int grmp_vector_destroy(grmp_vector *v)
{
    if (!v) return GRMP_OK;
    dfree(v->ctx, v->d);
    delete v;
    return GRMP_OK;
}

extern "C" int64_t grmp_vector_length(const grmp_vector *v) { return v ? v->n : -1; }

extern "C" int grmp_vector_upload(grmp_vector *v, const double *host)
{
    if (!v || (!host && v->n)) return GRMP_ERR_ARG;
    grmp_ctx *ctx = v->ctx;
    GRMP_HIP(ctx, hipMemcpyAsync(v->d, host, sizeof(double) * v->n, hipMemcpyHostToDevice, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GRMP_OK;
}

extern "C" int grmp_vector_download(const grmp_vector *v, double *host)
{
    if (!v || (!host && v->n)) return GRMP_ERR_ARG;
    grmp_ctx *ctx = v->ctx;
    GRMP_HIP(ctx, hipMemcpyAsync(host, v->d, sizeof(double) * v->n, hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GRMP_OK;
}

// page-locked host memory: device-to-host copies into it run as DMA without
// a staging copy, so a served solution lands on the host at link speed
extern "C" int grmp_host_alloc(grmp_ctx *ctx, size_t bytes, void **out)
{
    GRMP_CHECK(ctx, ctx && out, GRMP_ERR_ARG, "grmp_host_alloc: null argument");
    *out = nullptr;
    GRMP_HIP(ctx, hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return GRMP_OK;
}

extern "C" int grmp_host_free(void *p)
{
    if (p && hipHostFree(p) != hipSuccess) return GRMP_ERR_HIP;
    return GRMP_OK;
}

extern "C" int grmp_vector_fill(grmp_vector *v, double value)
{
    if (!v) return GRMP_ERR_ARG;
    if (v->n == 0) return GRMP_OK;
    hipLaunchKernelGGL(k_fill, dim3(nblk(v->n, 256)), dim3(256), 0, v->ctx->stream, v->d, v->n, value);
    GRMP_HIP(v->ctx, hipGetLastError());
    return GRMP_OK;
}

__global__ void k_scatter(double *v, int64_t n, const int64_t *idx, const double *val, int64_t cnt)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    int64_t i = idx[k] - 1;
    if (i >= 0 && i < n) v[i] = val[k];
}

extern "C" int grmp_vector_scatter(grmp_vector *v, int64_t count, const int64_t *idx, const double *values)
{
    if (!v || count < 0 || (count && (!idx || !values))) return GRMP_ERR_ARG;
    grmp_ctx *ctx = v->ctx;
    if (count == 0) return GRMP_OK;
    for (int64_t k = 0; k < count; ++k)
        GRMP_CHECK(ctx, idx[k] >= 1 && idx[k] <= v->n, GRMP_ERR_ARG, "grmp_vector_scatter: index %lld out of range",
                   (long long)idx[k]);
    std::lock_guard<std::mutex> g(ctx->mu);
    int64_t *di = nullptr;
    double *dv = nullptr;
    GRMP_TRY(dalloc(ctx, &di, count));
    GRMP_TRY(dalloc(ctx, &dv, count));
    GRMP_HIP(ctx, hipMemcpyAsync(di, idx, sizeof(int64_t) * count, hipMemcpyHostToDevice, ctx->stream));
    GRMP_HIP(ctx, hipMemcpyAsync(dv, values, sizeof(double) * count, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_scatter, dim3(nblk(count, 256)), dim3(256), 0, ctx->stream, v->d, v->n, di, dv, count);
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, di);
    dfree(ctx, dv);
    return GRMP_OK;
}

extern "C" int grmp_vector_lognormal(grmp_vector *v, uint64_t seed, double mu, double sigma)
{
    if (!v) return GRMP_ERR_ARG;
    if (v->n == 0) return GRMP_OK;
    hipLaunchKernelGGL(k_lognormal, dim3(nblk(v->n, 256)), dim3(256), 0, v->ctx->stream, v->d, v->n,
                       (uint32_t)seed, (uint32_t)(seed >> 32), mu, sigma);
    GRMP_HIP(v->ctx, hipGetLastError());
    return GRMP_OK;
}

// ----------------------------------------------------------------- meshes
__global__ void k_aos_to_soa_f64(const double *in, double *out, int64_t n, int dim)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < dim; ++d) out[d * n + i] = in[i * dim + d];
}

__global__ void k_aos_to_soa_i32_base1(const int32_t *in, int32_t *out, int64_t n, int nloc, int *bad,
                                       int64_t limit)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < nloc; ++d) {
        int32_t v = in[i * nloc + d] - 1;
        if (v < 0 || v >= limit) { *bad = 1; v = 0; }
        out[d * n + i] = v;
    }
}

// CellFaceSigns of the unit-square triangulation: +1 where the cell is the
// first to meet the face (its local orientation is the face's), else -1.
__global__ void k_unitsquare(int N, double *coords, int32_t *cells, int8_t *fsign)
{
    int64_t n1 = N + 1, nn = n1 * n1, nc = 2ll * N * N;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nn) {
        int64_t j = t / n1, i = t % n1;
        coords[t] = (double)i / (double)N;
        coords[nn + t] = (double)j / (double)N;
    }
    if (t < nc) {
        int64_t sq = t >> 1, j = sq / N, i = sq % N;
        int32_t v00 = (int32_t)(j * n1 + i), v10 = v00 + 1, v01 = (int32_t)(v00 + n1), v11 = v01 + 1;
        bool upper = t & 1;
        cells[t] = v00;
        cells[nc + t] = upper ? v11 : v10;
        cells[2 * nc + t] = upper ? v01 : v11;
        if (!upper) {
            fsign[t] = j == 0 ? 1 : -1;
            fsign[nc + t] = 1;
            fsign[2 * nc + t] = 1;
        } else {
            fsign[t] = -1;
            fsign[nc + t] = 1;
            fsign[2 * nc + t] = i == 0 ? 1 : -1;
        }
    }
}

__global__ void k_unitcube(int n, double jitter, uint32_t k0, uint32_t k1, double *coords, int32_t *cells)
{
    int64_t n1 = n + 1, nn = n1 * n1 * n1, nc = 6ll * n * n * n;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double h = 1.0 / (double)n;
    if (t < nn) {
        int64_t idx[3] = {t % n1, (t / n1) % n1, t / (n1 * n1)};
        uint32_t c[4] = {(uint32_t)t, (uint32_t)(t >> 32), 0x6a697474u, 0u};
        philox4x32(c, k0, k1);
        for (int d = 0; d < 3; ++d) {
            double x = (double)idx[d] / (double)n;
            if (idx[d] > 0 && idx[d] < n) x += jitter * h * (2.0 * u01(c[d]) - 1.0);
            coords[d * nn + t] = x;
        }
    }
    if (t < nc) {
        int64_t cube = t / 6;
        int p = (int)(t % 6);
        int64_t i = cube % n, j = (cube / n) % n, k = cube / ((int64_t)n * n);
        // Kuhn split: path 000 -> e_a -> e_a+e_b -> 111 for permutation p of (x,y,z)
        static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        int64_t pos[3] = {i, j, k};
        for (int v = 0; v < 4; ++v) {
            if (v > 0) pos[perm[p][v - 1]] += 1;
            cells[v * nc + t] = (int32_t)((pos[2] * n1 + pos[1]) * n1 + pos[0]);
        }
    }
}

// ------------------------------------------------- faces of triangle meshes
// Local face f of a cell joins its local nodes f and (f+1) % 3 (the order of
// the reference's Triangle2D face nodes).  Faces are numbered by first
// appearance, cell by cell, local face by local face; the face's global
// orientation is the one of its first appearance, and CellFaceSigns is +1
// where a cell traverses the face that way, -1 otherwise.
__global__ void k_face_keys(const int32_t *__restrict__ cells, int64_t nc, int64_t nn, uint64_t *__restrict__ key,
                            int32_t *__restrict__ app)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nc) return;
    const int64_t c = t / 3;
    const int f = (int)(t - 3 * c);
    const int64_t a = cells[f * nc + c], b = cells[((f + 1) % 3) * nc + c];
    key[t] = (uint64_t)(a < b ? a : b) * (uint64_t)nn + (uint64_t)(a < b ? b : a);
    app[t] = (int32_t)t;
}

// heads of equal-key segments (the sort is stable, so a head carries the
// segment's first appearance): mark[first appearance] = 1
__global__ void k_face_heads(const uint64_t *__restrict__ key, const int32_t *__restrict__ app, int64_t m,
                             int32_t *__restrict__ mark)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (i == 0 || key[i] != key[i - 1]) mark[app[i]] = 1;
}

// every appearance gets its segment's face id (rank of the segment's first
// appearance), its CellFaceSign and, for one-cell segments, the boundary flag
__global__ void k_face_assign(const uint64_t *__restrict__ key, const int32_t *__restrict__ app,
                              const int32_t *__restrict__ rank, const int32_t *__restrict__ cells, int64_t nc, int64_t m,
                              int32_t *__restrict__ cellfaces, int8_t *__restrict__ fsign, uint8_t *__restrict__ bface,
                              int *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t h = i;
    while (h > 0 && key[h - 1] == key[i]) --h;  // segments hold two entries on a conforming mesh
    if (i - h > 1) atomicOr(bad, 1);           // a face shared by three or more cells
    const int32_t first = app[h];
    const int32_t id = rank[first];
    const int32_t me = app[i];
    const int64_t c = me / 3, c0 = first / 3;
    const int f = me - 3 * (int)c, f0 = first - 3 * (int)c0;
    cellfaces[f * nc + c] = id;
    const int32_t a = cells[f * nc + c], a0 = cells[f0 * nc + c0];
    fsign[f * nc + c] = (int8_t)(a == a0 ? 1 : -1);
    const bool single = (i == h) && (i + 1 >= m || key[i + 1] != key[i]);
    if (single) bface[id] = 1;
}

int mesh_build_faces(grmp_ctx *ctx, grmp_mesh *m)
{
    const int64_t nc = m->ncells, M = 3 * nc;
    GRMP_CHECK(ctx, m->dim == 2 && m->nloc == 3 && M < (1ll << 31), GRMP_ERR_ARG, "mesh faces: triangle meshes only");
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int32_t *a0 = nullptr, *a1 = nullptr, *mark = nullptr, *rank = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &k0, M)) || (rc = dalloc(ctx, &k1, M)) || (rc = dalloc(ctx, &a0, M)) ||
            (rc = dalloc(ctx, &a1, M)) || (rc = dalloc(ctx, &mark, M)) || (rc = dalloc(ctx, &rank, M + 1)))
            break;
        hipLaunchKernelGGL(k_face_keys, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, m->cells, nc, m->nnodes, k0, a0);
        // keys are below nnodes^2: sort only the bits they use
        int bits = 1;
        while (bits < 64 && ((uint64_t)1 << bits) < (uint64_t)m->nnodes * (uint64_t)m->nnodes) ++bits;
        size_t tmp = 0;
        if (rocprim::radix_sort_pairs(nullptr, tmp, k0, k1, a0, a1, (size_t)M, 0u, (unsigned)bits, ctx->stream) !=
            hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "mesh faces: sort setup");
            break;
        }
        if ((rc = ensure_scratch(ctx, tmp))) break;
        if (rocprim::radix_sort_pairs(ctx->scratch, tmp, k0, k1, a0, a1, (size_t)M, 0u, (unsigned)bits,
                                      ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "mesh faces: sort");
            break;
        }
        if (hipMemsetAsync(mark, 0, sizeof(int32_t) * M, ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "mesh faces: memset");
            break;
        }
        hipLaunchKernelGGL(k_face_heads, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, k1, a1, M, mark);
        if ((rc = exclusive_scan_i32(ctx, mark, rank, M))) break;
        int32_t nf = 0;
        if (hipMemcpyAsync(&nf, rank + M, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "mesh faces: count");
            break;
        }
        m->nfaces = nf;
        if ((rc = dalloc(ctx, &m->cellfaces, M)) || (rc = dalloc(ctx, &m->bface, nf)) ||
            (!m->fsign && (rc = dalloc(ctx, &m->fsign, M))))
            break;
        if (hipMemsetAsync(m->bface, 0, nf, ctx->stream) != hipSuccess ||
            hipMemsetAsync(ctx->dflag, 0, sizeof(int), ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "mesh faces: memset");
            break;
        }
        hipLaunchKernelGGL(k_face_assign, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, k1, a1, rank, m->cells, nc, M,
                           m->cellfaces, m->fsign, m->bface, ctx->dflag);
        int bad = 0;
        if (hipMemcpyAsync(&bad, ctx->dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "mesh faces: assign");
            break;
        }
        if (bad) rc = grmp_fail(ctx, GRMP_ERR_ARG, "mesh faces: a face is shared by more than two cells");
    } while (0);
    dfree(ctx, k0);
    dfree(ctx, k1);
    dfree(ctx, a0);
    dfree(ctx, a1);
    dfree(ctx, mark);
    dfree(ctx, rank);
    return rc;
}

static grmp_mesh *mesh_new(grmp_ctx *ctx, int dim, int64_t nnodes, int64_t ncells, int nloc)
{
    grmp_mesh *m = new grmp_mesh();
    m->ctx = ctx;
    m->dim = dim;
    m->nnodes = nnodes;
    m->ncells = ncells;
    m->nloc = nloc;
    m->kind = MESH_GENERIC;
    m->N = 0;
    m->coords = nullptr;
    m->cells = nullptr;
    m->fsign = nullptr;
    m->cellfaces = nullptr;
    m->bface = nullptr;
    m->nfaces = 0;
    return m;
}

extern "C" int grmp_mesh_destroy(grmp_mesh *m)
{
    if (!m) return GRMP_OK;
    dfree(m->ctx, m->coords);
    dfree(m->ctx, m->cells);
    dfree(m->ctx, m->fsign);
    dfree(m->ctx, m->cellfaces);
    dfree(m->ctx, m->bface);
    delete m;
    return GRMP_OK;
}

extern "C" int grmp_mesh_upload(grmp_ctx *ctx, int dim, int64_t nnodes, const double *coords, int64_t ncells,
                                int nloc, const int32_t *cellnodes, grmp_mesh **out)
{
    GRMP_CHECK(ctx, ctx && out && coords && cellnodes, GRMP_ERR_ARG, "grmp_mesh_upload: null argument");
    GRMP_CHECK(ctx, (dim == 2 && nloc == 3) || (dim == 3 && nloc == 4), GRMP_ERR_UNSUPPORTED,
               "grmp_mesh_upload: only Triangle2D (dim 2, 3 nodes) and Tetrahedron3D (dim 3, 4 nodes)");
    GRMP_CHECK(ctx, nnodes > 0 && ncells > 0 && nnodes < (1ll << 31) && ncells * nloc < (1ll << 31),
               GRMP_ERR_ARG, "grmp_mesh_upload: sizes out of range");
    std::lock_guard<std::mutex> g(ctx->mu);
    grmp_mesh *m = mesh_new(ctx, dim, nnodes, ncells, nloc);
    double *tc = nullptr;
    int32_t *tn = nullptr;
    int rc = GRMP_OK;
    if ((rc = dalloc(ctx, &m->coords, dim * nnodes)) || (rc = dalloc(ctx, &m->cells, nloc * ncells)) ||
        (rc = dalloc(ctx, &tc, dim * nnodes)) || (rc = dalloc(ctx, &tn, nloc * ncells))) {
        dfree(ctx, tc); dfree(ctx, tn);
        grmp_mesh_destroy(m);
        return rc;
    }
    (void)hipMemcpyAsync(tc, coords, sizeof(double) * dim * nnodes, hipMemcpyHostToDevice, ctx->stream);
    (void)hipMemcpyAsync(tn, cellnodes, sizeof(int32_t) * nloc * ncells, hipMemcpyHostToDevice, ctx->stream);
    (void)hipMemsetAsync(ctx->dflag, 0, sizeof(int), ctx->stream);
    hipLaunchKernelGGL(k_aos_to_soa_f64, dim3(nblk(nnodes, 256)), dim3(256), 0, ctx->stream, tc, m->coords,
                       nnodes, dim);
    hipLaunchKernelGGL(k_aos_to_soa_i32_base1, dim3(nblk(ncells, 256)), dim3(256), 0, ctx->stream, tn,
                       m->cells, ncells, nloc, ctx->dflag, nnodes);
    int bad = 0;
    (void)hipMemcpyAsync(&bad, ctx->dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, tc);
    dfree(ctx, tn);
    if (e != hipSuccess) {
        grmp_mesh_destroy(m);
        return grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mesh_upload: %s", hipGetErrorString(e));
    }
    if (bad) {
        grmp_mesh_destroy(m);
        return grmp_fail(ctx, GRMP_ERR_ARG, "grmp_mesh_upload: cell node index outside 1..nnodes");
    }
    if (dim == 2 && (rc = mesh_build_faces(ctx, m))) {
        grmp_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return GRMP_OK;
}

extern "C" int grmp_mesh_unitsquare(grmp_ctx *ctx, int N, grmp_mesh **out)
{
    GRMP_CHECK(ctx, ctx && out && N >= 1 && N <= 16384, GRMP_ERR_ARG, "grmp_mesh_unitsquare: N out of range");
    std::lock_guard<std::mutex> g(ctx->mu);
    int64_t nn = (int64_t)(N + 1) * (N + 1), nc = 2ll * N * N;
    grmp_mesh *m = mesh_new(ctx, 2, nn, nc, 3);
    int rc;
    if ((rc = dalloc(ctx, &m->coords, 2 * nn)) || (rc = dalloc(ctx, &m->cells, 3 * nc)) ||
        (rc = dalloc(ctx, &m->fsign, 3 * nc))) {
        grmp_mesh_destroy(m);
        return rc;
    }
    m->kind = MESH_UNITSQUARE;
    m->N = N;
    int64_t tot = nn > nc ? nn : nc;
    hipLaunchKernelGGL(k_unitsquare, dim3(nblk(tot, 256)), dim3(256), 0, ctx->stream, N, m->coords, m->cells, m->fsign);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { grmp_mesh_destroy(m); return grmp_fail(ctx, GRMP_ERR_HIP, "k_unitsquare: %s", hipGetErrorString(e)); }
    *out = m;
    return GRMP_OK;
}

extern "C" int grmp_mesh_unitcube(grmp_ctx *ctx, int n, double jitter, uint64_t seed, grmp_mesh **out)
{
    GRMP_CHECK(ctx, ctx && out && n >= 1 && n <= 700 && jitter >= 0.0 && jitter < 0.25, GRMP_ERR_ARG,
               "grmp_mesh_unitcube: n or jitter out of range");
    std::lock_guard<std::mutex> g(ctx->mu);
    int64_t nn = (int64_t)(n + 1) * (n + 1) * (n + 1), nc = 6ll * n * n * n;
    grmp_mesh *m = mesh_new(ctx, 3, nn, nc, 4);
    int rc;
    if ((rc = dalloc(ctx, &m->coords, 3 * nn)) || (rc = dalloc(ctx, &m->cells, 4 * nc))) {
        grmp_mesh_destroy(m);
        return rc;
    }
    m->kind = MESH_UNITCUBE;
    m->N = n;
    int64_t tot = nn > nc ? nn : nc;
    hipLaunchKernelGGL(k_unitcube, dim3(nblk(tot, 256)), dim3(256), 0, ctx->stream, n, jitter, (uint32_t)seed,
                       (uint32_t)(seed >> 32), m->coords, m->cells);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { grmp_mesh_destroy(m); return grmp_fail(ctx, GRMP_ERR_HIP, "k_unitcube: %s", hipGetErrorString(e)); }
    *out = m;
    return GRMP_OK;
}

extern "C" int grmp_mesh_info(const grmp_mesh *m, int *dim, int64_t *nnodes, int64_t *ncells)
{
    if (!m) return GRMP_ERR_ARG;
    if (dim) *dim = m->dim;
    if (nnodes) *nnodes = m->nnodes;
    if (ncells) *ncells = m->ncells;
    return GRMP_OK;
}

__global__ void k_soa_to_aos_f64(const double *in, double *out, int64_t n, int dim)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < dim; ++d) out[i * dim + d] = in[d * n + i];
}

__global__ void k_soa_to_aos_i32_base1(const int32_t *in, int32_t *out, int64_t n, int nloc)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < nloc; ++d) out[i * nloc + d] = in[d * n + i] + 1;
}

extern "C" int grmp_mesh_download(const grmp_mesh *m, double *coords, int32_t *cellnodes)
{
    if (!m) return GRMP_ERR_ARG;
    grmp_ctx *ctx = m->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    double *tc = nullptr;
    int32_t *tn = nullptr;
    GRMP_TRY(dalloc(ctx, &tc, m->dim * m->nnodes));
    GRMP_TRY(dalloc(ctx, &tn, m->nloc * m->ncells));
    hipLaunchKernelGGL(k_soa_to_aos_f64, dim3(nblk(m->nnodes, 256)), dim3(256), 0, ctx->stream, m->coords, tc,
                       m->nnodes, m->dim);
    hipLaunchKernelGGL(k_soa_to_aos_i32_base1, dim3(nblk(m->ncells, 256)), dim3(256), 0, ctx->stream, m->cells,
                       tn, m->ncells, m->nloc);
    if (coords) (void)hipMemcpyAsync(coords, tc, sizeof(double) * m->dim * m->nnodes, hipMemcpyDeviceToHost, ctx->stream);
    if (cellnodes) (void)hipMemcpyAsync(cellnodes, tn, sizeof(int32_t) * m->nloc * m->ncells, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, tc);
    dfree(ctx, tn);
    GRMP_HIP(ctx, e);
    return GRMP_OK;
}

// ----------------------------------------------------------------- spaces

#include "grmp_square.h"

__global__ void k_celldofs_p1(const int32_t *cells, int32_t *dofs, int64_t nc, int nv, int ncomp, int64_t nnodes)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int nloc = nv * ncomp;
    for (int k = 0; k < ncomp; ++k)
        for (int a = 0; a < nv; ++a) dofs[(int64_t)(k * nv + a) * nc + c] = cells[(int64_t)a * nc + c] + (int32_t)(k * nnodes);
    (void)nloc;
}

__global__ void k_celldofs_p2_square(const int32_t *cells, int32_t *dofs, int64_t nc, int N, int ncomp,
                                     int64_t nnodes, int64_t nfaces)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int64_t f[3];
    sq_cellfaces(N, c, f);
    int64_t coff = nnodes + nfaces;
    for (int k = 0; k < ncomp; ++k) {
        for (int a = 0; a < 3; ++a) dofs[(int64_t)(k * 6 + a) * nc + c] = (int32_t)(cells[(int64_t)a * nc + c] + k * coff);
        for (int a = 0; a < 3; ++a) dofs[(int64_t)(k * 6 + 3 + a) * nc + c] = (int32_t)(nnodes + f[a] + k * coff);
    }
}

// H1BR{2} "N1f1": node dofs of both components (offset nnodes), then one
// dof per face (offset 2 nnodes), faces numbered by first appearance
__global__ void k_celldofs_br_square(const int32_t *cells, int32_t *dofs, int64_t nc, int N, int64_t nnodes)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int64_t f[3];
    sq_cellfaces(N, c, f);
    for (int a = 0; a < 3; ++a) {
        dofs[(int64_t)a * nc + c] = cells[(int64_t)a * nc + c];
        dofs[(int64_t)(3 + a) * nc + c] = (int32_t)(cells[(int64_t)a * nc + c] + nnodes);
        dofs[(int64_t)(6 + a) * nc + c] = (int32_t)(2 * nnodes + f[a]);
    }
}

// H1P2 / H1BR{2} dofmaps from enumerated faces (uploaded triangle meshes)
__global__ void k_celldofs_p2_faces(const int32_t *cells, const int32_t *cellfaces, int32_t *dofs, int64_t nc, int ncomp,
                                    int64_t nnodes, int64_t nfaces)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int64_t coff = nnodes + nfaces;
    for (int k = 0; k < ncomp; ++k) {
        for (int a = 0; a < 3; ++a) dofs[(int64_t)(k * 6 + a) * nc + c] = (int32_t)(cells[(int64_t)a * nc + c] + k * coff);
        for (int a = 0; a < 3; ++a)
            dofs[(int64_t)(k * 6 + 3 + a) * nc + c] = (int32_t)(nnodes + cellfaces[(int64_t)a * nc + c] + k * coff);
    }
}

__global__ void k_celldofs_br_faces(const int32_t *cells, const int32_t *cellfaces, int32_t *dofs, int64_t nc,
                                    int64_t nnodes)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int a = 0; a < 3; ++a) {
        dofs[(int64_t)a * nc + c] = cells[(int64_t)a * nc + c];
        dofs[(int64_t)(3 + a) * nc + c] = (int32_t)(cells[(int64_t)a * nc + c] + nnodes);
        dofs[(int64_t)(6 + a) * nc + c] = (int32_t)(2 * nnodes + cellfaces[(int64_t)a * nc + c]);
    }
}

// L2P0 "I1": the cell's own dof
__global__ void k_celldofs_p0(int32_t *dofs, int64_t nc)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) dofs[c] = (int32_t)c;
}

__global__ void k_inc_count(const int32_t *dofs, int64_t nc, int nloc, int32_t *cnt)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * nloc) return;
    atomicAdd(&cnt[dofs[t]], 1);
}

__global__ void k_inc_fill(const int32_t *dofs, int64_t nc, int nloc, const int32_t *ptr, int32_t *cur, int32_t *inc)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * nloc) return;
    int64_t a = t / nc, c = t % nc;
    int32_t d = dofs[t];
    int32_t p = atomicAdd(&cur[d], 1);
    inc[ptr[d] + p] = (int32_t)(c * nloc + a);
}

// sorts each dof's incidence list by (cell, local index): the order in which
// the reference's element loop visits them
__global__ void k_inc_sort(const int32_t *ptr, int32_t *inc, int64_t ndofs, int *maxinc)
{
    int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndofs) return;
    int32_t b = ptr[d], e = ptr[d + 1];
    for (int32_t i = b + 1; i < e; ++i) {
        int32_t v = inc[i];
        int32_t j = i - 1;
        while (j >= b && inc[j] > v) { inc[j + 1] = inc[j]; --j; }
        inc[j + 1] = v;
    }
    atomicMax(maxinc, e - b);
}

int build_incidence(grmp_ctx *ctx, grmp_space *s)
{
    int64_t nc = s->mesh->ncells, tot = nc * s->nloc;
    int32_t *cnt = nullptr;
    GRMP_TRY(dalloc(ctx, &s->inc_ptr, s->ndofs + 1));
    GRMP_TRY(dalloc(ctx, &s->inc, tot));
    GRMP_TRY(dalloc(ctx, &cnt, s->ndofs));
    GRMP_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int32_t) * s->ndofs, ctx->stream));
    hipLaunchKernelGGL(k_inc_count, dim3(nblk(tot, 256)), dim3(256), 0, ctx->stream, s->celldofs, nc, s->nloc, cnt);
    GRMP_TRY(exclusive_scan_i32(ctx, cnt, s->inc_ptr, s->ndofs));
    GRMP_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int32_t) * s->ndofs, ctx->stream));
    // celldofs is SoA [nloc][nc]: entry t = a*nc + c
    hipLaunchKernelGGL(k_inc_fill, dim3(nblk(tot, 256)), dim3(256), 0, ctx->stream, s->celldofs, nc, s->nloc,
                       s->inc_ptr, cnt, s->inc);
    GRMP_HIP(ctx, hipMemsetAsync(ctx->dflag, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_inc_sort, dim3(nblk(s->ndofs, 256)), dim3(256), 0, ctx->stream, s->inc_ptr, s->inc,
                       s->ndofs, ctx->dflag);
    int mx = 0;
    GRMP_HIP(ctx, hipMemcpyAsync(&mx, ctx->dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, cnt);
    s->max_inc = mx;
    return GRMP_OK;
}

static int fe_nloc(int fe, int dim)
{
    if (fe == GRMP_FE_H1P1) return dim + 1;
    if (fe == GRMP_FE_H1P2) return dim == 2 ? 6 : 10;
    if (fe == GRMP_FE_L2P0) return 1;
    if (fe == GRMP_FE_H1BR) return 9;
    return 0;
}

extern "C" int grmp_space_destroy(grmp_space *s)
{
    if (!s) return GRMP_OK;
    dfree(s->ctx, s->celldofs);
    dfree(s->ctx, s->inc_ptr);
    dfree(s->ctx, s->inc);
    delete s;
    return GRMP_OK;
}

extern "C" int grmp_space_upload(grmp_ctx *ctx, const grmp_mesh *mesh, int fe, int ncomp, int64_t ndofs,
                                 int ndofs_loc, const int32_t *celldofs, grmp_space **out)
{
    GRMP_CHECK(ctx, ctx && mesh && out && celldofs, GRMP_ERR_ARG, "grmp_space_upload: null argument");
    GRMP_CHECK(ctx, ncomp >= 1 && ncomp <= 3 && ndofs > 0 && ndofs < (1ll << 31), GRMP_ERR_ARG,
               "grmp_space_upload: bad ncomp/ndofs");
    int base = fe_nloc(fe, mesh->dim);
    GRMP_CHECK(ctx, base > 0, GRMP_ERR_UNSUPPORTED, "grmp_space_upload: unsupported FE type %d", fe);
    int expect = (fe == GRMP_FE_H1BR) ? 9 : base * ncomp;
    GRMP_CHECK(ctx, ndofs_loc == expect, GRMP_ERR_ARG, "grmp_space_upload: ndofs_loc %d != %d", ndofs_loc, expect);
    std::lock_guard<std::mutex> g(ctx->mu);
    grmp_space *s = new grmp_space();
    s->ctx = ctx;
    s->mesh = mesh;
    s->fe = fe;
    s->ncomp = ncomp;
    s->ndofs = ndofs;
    s->nloc = ndofs_loc;
    s->nfaces = 0;
    int64_t nc = mesh->ncells;
    int32_t *tmp = nullptr;
    int rc;
    if ((rc = dalloc(ctx, &s->celldofs, nc * ndofs_loc)) || (rc = dalloc(ctx, &tmp, nc * ndofs_loc))) {
        dfree(ctx, tmp);
        grmp_space_destroy(s);
        return rc;
    }
    (void)hipMemcpyAsync(tmp, celldofs, sizeof(int32_t) * nc * ndofs_loc, hipMemcpyHostToDevice, ctx->stream);
    (void)hipMemsetAsync(ctx->dflag, 0, sizeof(int), ctx->stream);
    hipLaunchKernelGGL(k_aos_to_soa_i32_base1, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, tmp, s->celldofs,
                       nc, ndofs_loc, ctx->dflag, ndofs);
    int bad = 0;
    (void)hipMemcpyAsync(&bad, ctx->dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, tmp);
    if (e != hipSuccess || bad) {
        grmp_space_destroy(s);
        return grmp_fail(ctx, bad ? GRMP_ERR_ARG : GRMP_ERR_HIP, "grmp_space_upload: %s",
                         bad ? "dof index outside 1..ndofs" : hipGetErrorString(e));
    }
    if ((rc = build_incidence(ctx, s))) { grmp_space_destroy(s); return rc; }
    *out = s;
    return GRMP_OK;
}

extern "C" int grmp_space_create(grmp_ctx *ctx, const grmp_mesh *mesh, int fe, int ncomp, grmp_space **out)
{
    GRMP_CHECK(ctx, ctx && mesh && out && ncomp >= 1 && ncomp <= 3, GRMP_ERR_ARG, "grmp_space_create: bad arguments");
    const bool faces = mesh->kind == MESH_UNITSQUARE || mesh->cellfaces;
    GRMP_CHECK(ctx, fe == GRMP_FE_H1P1 || fe == GRMP_FE_L2P0 || ((fe == GRMP_FE_H1P2 || fe == GRMP_FE_H1BR) && faces),
               GRMP_ERR_UNSUPPORTED,
               "grmp_space_create: device dofmaps exist for H1P1/L2P0 on any mesh and H1P2/H1BR on triangle "
               "meshes; upload FES[CellDofs] with grmp_space_upload otherwise");
    GRMP_CHECK(ctx, fe != GRMP_FE_H1BR || ncomp == 2, GRMP_ERR_ARG, "grmp_space_create: H1BR{2} has 2 components");
    GRMP_CHECK(ctx, fe != GRMP_FE_L2P0 || ncomp == 1, GRMP_ERR_UNSUPPORTED, "grmp_space_create: L2P0 with 1 component");
    std::lock_guard<std::mutex> g(ctx->mu);
    grmp_space *s = new grmp_space();
    s->ctx = ctx;
    s->mesh = mesh;
    s->fe = fe;
    s->ncomp = ncomp;
    int nv = mesh->nloc;
    int64_t nc = mesh->ncells;
    if (fe == GRMP_FE_H1P1) {
        s->nloc = nv * ncomp;
        s->ndofs = mesh->nnodes * ncomp;
        s->nfaces = 0;
    } else if (fe == GRMP_FE_L2P0) {
        s->nloc = 1;
        s->ndofs = nc;
        s->nfaces = 0;
    } else if (fe == GRMP_FE_H1BR) {
        int64_t N = mesh->N;
        s->nfaces = mesh->cellfaces ? mesh->nfaces : 3 * N * N + 2 * N;
        s->nloc = 9;
        s->ndofs = 2 * mesh->nnodes + s->nfaces;
    } else {
        int64_t N = mesh->N;
        s->nfaces = mesh->cellfaces ? mesh->nfaces : 3 * N * N + 2 * N;
        s->nloc = 6 * ncomp;
        s->ndofs = (mesh->nnodes + s->nfaces) * ncomp;
    }
    int rc;
    if ((rc = dalloc(ctx, &s->celldofs, nc * s->nloc))) { grmp_space_destroy(s); return rc; }
    if (fe == GRMP_FE_H1P1)
        hipLaunchKernelGGL(k_celldofs_p1, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, mesh->cells, s->celldofs, nc,
                           nv, ncomp, mesh->nnodes);
    else if (fe == GRMP_FE_L2P0)
        hipLaunchKernelGGL(k_celldofs_p0, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, s->celldofs, nc);
    else if (fe == GRMP_FE_H1BR && mesh->cellfaces)
        hipLaunchKernelGGL(k_celldofs_br_faces, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, mesh->cells,
                           mesh->cellfaces, s->celldofs, nc, mesh->nnodes);
    else if (fe == GRMP_FE_H1BR)
        hipLaunchKernelGGL(k_celldofs_br_square, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, mesh->cells,
                           s->celldofs, nc, mesh->N, mesh->nnodes);
    else if (mesh->cellfaces)
        hipLaunchKernelGGL(k_celldofs_p2_faces, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, mesh->cells,
                           mesh->cellfaces, s->celldofs, nc, ncomp, mesh->nnodes, s->nfaces);
    else
        hipLaunchKernelGGL(k_celldofs_p2_square, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, mesh->cells,
                           s->celldofs, nc, mesh->N, ncomp, mesh->nnodes, s->nfaces);
    if ((rc = build_incidence(ctx, s))) { grmp_space_destroy(s); return rc; }
    *out = s;
    return GRMP_OK;
}

extern "C" int grmp_space_info(const grmp_space *s, int64_t *ndofs, int *ndofs_loc)
{
    if (!s) return GRMP_ERR_ARG;
    if (ndofs) *ndofs = s->ndofs;
    if (ndofs_loc) *ndofs_loc = s->nloc;
    return GRMP_OK;
}

extern "C" int grmp_space_download_celldofs(const grmp_space *s, int32_t *celldofs)
{
    if (!s || !celldofs) return GRMP_ERR_ARG;
    grmp_ctx *ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    int32_t *t = nullptr;
    int64_t nc = s->mesh->ncells;
    GRMP_TRY(dalloc(ctx, &t, nc * s->nloc));
    hipLaunchKernelGGL(k_soa_to_aos_i32_base1, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, s->celldofs, t, nc,
                       s->nloc);
    (void)hipMemcpyAsync(celldofs, t, sizeof(int32_t) * nc * s->nloc, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, t);
    GRMP_HIP(ctx, e);
    return GRMP_OK;
}

// boundary dofs of structured meshes, enumerated on the host (O(boundary))
// CellFaces / CellFaceSigns in the caller's layout ([ncells][3], faces 1-based)
__global__ void k_faces_out(const int32_t *__restrict__ cellfaces, const int8_t *__restrict__ fsign, int64_t nc, int N,
                            int32_t *__restrict__ cf, int8_t *__restrict__ cs)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int64_t f[3];
    if (cellfaces)
        for (int k = 0; k < 3; ++k) f[k] = cellfaces[k * nc + c];
    else
        sq_cellfaces(N, c, f);
    for (int k = 0; k < 3; ++k) {
        cf[3 * c + k] = (int32_t)(f[k] + 1);
        cs[3 * c + k] = fsign[k * nc + c];
    }
}

// boundary faces of the generated unit square: bottom, top, left, right
__global__ void k_bface_square(int N, uint8_t *__restrict__ bface)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * (int64_t)N) return;
    const int64_t side = t / N, i = t % N;
    const int64_t f = side == 0 ? sq_H(N, i, 0) : side == 1 ? sq_H(N, i, N) : side == 2 ? sq_V(N, 0, i) : sq_V(N, N, i);
    bface[f] = 1;
}

extern "C" int grmp_mesh_faces(const grmp_mesh *m, int64_t *nfaces, int32_t *cellfaces, int8_t *cellfacesigns,
                               uint8_t *bface)
{
    if (!m || !nfaces) return GRMP_ERR_ARG;
    grmp_ctx *ctx = m->ctx;
    const bool square = m->kind == MESH_UNITSQUARE;
    GRMP_CHECK(ctx, square || m->cellfaces, GRMP_ERR_UNSUPPORTED,
               "grmp_mesh_faces: face adjacency exists for triangle meshes");
    const int64_t nc = m->ncells, N = m->N;
    const int64_t nf = square ? 3 * N * N + 2 * N : m->nfaces;
    *nfaces = nf;
    std::lock_guard<std::mutex> g(ctx->mu);
    int32_t *cf = nullptr;
    int8_t *cs = nullptr;
    uint8_t *bf = nullptr;
    int rc = GRMP_OK;
    do {
        if (cellfaces || cellfacesigns) {
            if ((rc = dalloc(ctx, &cf, 3 * nc)) || (rc = dalloc(ctx, &cs, 3 * nc))) break;
            hipLaunchKernelGGL(k_faces_out, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream,
                               square ? nullptr : m->cellfaces, m->fsign, nc, (int)N, cf, cs);
            if ((cellfaces && hipMemcpyAsync(cellfaces, cf, sizeof(int32_t) * 3 * nc, hipMemcpyDeviceToHost,
                                             ctx->stream) != hipSuccess) ||
                (cellfacesigns &&
                 hipMemcpyAsync(cellfacesigns, cs, 3 * nc, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)) {
                rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mesh_faces: download");
                break;
            }
        }
        if (bface) {
            const uint8_t *src = m->bface;
            if (square) {
                if ((rc = dalloc(ctx, &bf, nf))) break;
                if (hipMemsetAsync(bf, 0, nf, ctx->stream) != hipSuccess) {
                    rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mesh_faces: memset");
                    break;
                }
                hipLaunchKernelGGL(k_bface_square, dim3(nblk(4 * N, 256)), dim3(256), 0, ctx->stream, (int)N, bf);
                src = bf;
            }
            if (hipMemcpyAsync(bface, src, nf, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) {
                rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mesh_faces: download");
                break;
            }
        }
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_mesh_faces: sync");
    } while (0);
    dfree(ctx, cf);
    dfree(ctx, cs);
    dfree(ctx, bf);
    return rc;
}

// BFaceDofs of enumerated boundary faces: every (cell, local face) on the
// boundary marks the face's two nodes in each component and, for P2 / BR, the
// face's own dof (src/dofmaps.jl BFaceDofs patterns "N1F1" / "N1i1")
__global__ void k_mark_bdofs(const int32_t *__restrict__ cells, const int32_t *__restrict__ cellfaces,
                             const uint8_t *__restrict__ bface, int64_t nc, int64_t nnodes, int fe, int ncomp,
                             int64_t coff, uint8_t *__restrict__ mark)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nc) return;
    const int64_t c = t / 3;
    const int f = (int)(t - 3 * c);
    const int32_t face = cellfaces[f * nc + c];
    if (!bface[face]) return;
    const int64_t a = cells[f * nc + c], b = cells[((f + 1) % 3) * nc + c];
    if (fe == GRMP_FE_H1BR) {
        mark[a] = mark[b] = mark[a + nnodes] = mark[b + nnodes] = 1;
        mark[2 * nnodes + face] = 1;
        return;
    }
    for (int k = 0; k < ncomp; ++k) {
        mark[a + k * coff] = mark[b + k * coff] = 1;
        if (fe == GRMP_FE_H1P2) mark[nnodes + face + k * coff] = 1;
    }
}

static int boundary_dofs_faces(grmp_ctx *ctx, const grmp_space *s, int64_t *count, int64_t *out)
{
    const grmp_mesh *m = s->mesh;
    const int64_t nc = m->ncells, nd = s->ndofs;
    uint8_t *mark = nullptr;
    int64_t *sel = nullptr, *nsel = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &mark, nd)) || (rc = dalloc(ctx, &sel, nd)) || (rc = dalloc(ctx, &nsel, 1))) break;
        if (hipMemsetAsync(mark, 0, nd, ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "boundary dofs: memset");
            break;
        }
        const int64_t coff = s->ndofs / s->ncomp;
        hipLaunchKernelGGL(k_mark_bdofs, dim3(nblk(3 * nc, 256)), dim3(256), 0, ctx->stream, m->cells, m->cellfaces,
                           m->bface, nc, m->nnodes, s->fe, s->ncomp, coff, mark);
        // stream compaction keeps index order, so the list comes out sorted
        rocprim::counting_iterator<int64_t> idx(1);  // 1-based dof numbers
        size_t tmp = 0;
        if (rocprim::select(nullptr, tmp, idx, mark, sel, nsel, (size_t)nd, ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "boundary dofs: select setup");
            break;
        }
        if ((rc = ensure_scratch(ctx, tmp))) break;
        if (rocprim::select(ctx->scratch, tmp, idx, mark, sel, nsel, (size_t)nd, ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "boundary dofs: select");
            break;
        }
        int64_t cnt = 0;
        if (hipMemcpyAsync(&cnt, nsel, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "boundary dofs: count");
            break;
        }
        *count = cnt;
        if (out && cnt &&
            (hipMemcpyAsync(out, sel, sizeof(int64_t) * cnt, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess))
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "boundary dofs: download");
    } while (0);
    dfree(ctx, mark);
    dfree(ctx, sel);
    dfree(ctx, nsel);
    return rc;
}

extern "C" int grmp_space_boundary_dofs(const grmp_space *s, int64_t *count, int64_t *out)
{
    if (!s || !count) return GRMP_ERR_ARG;
    const grmp_mesh *m = s->mesh;
    grmp_ctx *ctx = s->ctx;
    if (s->fe == GRMP_FE_L2P0) {
        *count = 0;  // no dofs on faces
        return GRMP_OK;
    }
    if (m->kind != MESH_UNITSQUARE && m->cellfaces) {
        std::lock_guard<std::mutex> g(ctx->mu);
        return boundary_dofs_faces(ctx, s, count, out);
    }
    GRMP_CHECK(ctx, m->kind == MESH_UNITSQUARE || m->kind == MESH_UNITCUBE, GRMP_ERR_UNSUPPORTED,
               "grmp_space_boundary_dofs: boundary enumeration exists for generated meshes and uploaded "
               "triangle meshes");
    std::vector<int64_t> b;
    int64_t N = m->N, n1 = N + 1;
    if (m->kind == MESH_UNITSQUARE && s->fe == GRMP_FE_H1BR) {
        // BFaceDofs "N1i1": both components of boundary nodes, and the bubble of each boundary face
        std::vector<int64_t> nodes;
        for (int64_t j = 0; j <= N; ++j)
            for (int64_t i = 0; i <= N; ++i)
                if (i == 0 || j == 0 || i == N || j == N) nodes.push_back(j * n1 + i);
        for (int64_t v : nodes) {
            b.push_back(v);
            b.push_back(v + m->nnodes);
        }
        for (int64_t i = 0; i < N; ++i) {
            b.push_back(2 * m->nnodes + sq_H(N, i, 0));
            b.push_back(2 * m->nnodes + sq_H(N, i, N));
        }
        for (int64_t j = 0; j < N; ++j) {
            b.push_back(2 * m->nnodes + sq_V(N, 0, j));
            b.push_back(2 * m->nnodes + sq_V(N, N, j));
        }
        std::sort(b.begin(), b.end());
        *count = (int64_t)b.size();
        if (out)
            for (size_t q = 0; q < b.size(); ++q) out[q] = b[q] + 1;
        return GRMP_OK;
    }
    if (m->kind == MESH_UNITSQUARE) {
        for (int64_t j = 0; j <= N; ++j)
            for (int64_t i = 0; i <= N; ++i)
                if (i == 0 || j == 0 || i == N || j == N) b.push_back(j * n1 + i);
        if (s->fe == GRMP_FE_H1P2) {
            for (int64_t i = 0; i < N; ++i) {
                b.push_back(m->nnodes + sq_H(N, i, 0));
                b.push_back(m->nnodes + sq_H(N, i, N));
            }
            for (int64_t j = 0; j < N; ++j) {
                b.push_back(m->nnodes + sq_V(N, 0, j));
                b.push_back(m->nnodes + sq_V(N, N, j));
            }
        }
    } else {
        for (int64_t k = 0; k <= N; ++k)
            for (int64_t j = 0; j <= N; ++j)
                for (int64_t i = 0; i <= N; ++i)
                    if (i == 0 || j == 0 || k == 0 || i == N || j == N || k == N) b.push_back((k * n1 + j) * n1 + i);
    }
    int64_t per = s->ndofs / s->ncomp;
    size_t nb = b.size();
    for (int c = 1; c < s->ncomp; ++c)
        for (size_t q = 0; q < nb; ++q) b.push_back(b[q] + c * per);
    std::sort(b.begin(), b.end());
    *count = (int64_t)b.size();
    if (out)
        for (size_t q = 0; q < b.size(); ++q) out[q] = b[q] + 1;
    return GRMP_OK;
}

extern "C" int grmp_mesh_cellvolumes(const grmp_mesh *m, double *out);
