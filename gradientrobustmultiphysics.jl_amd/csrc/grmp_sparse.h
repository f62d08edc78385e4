// Device helpers shared by the Krylov solvers: fixed-order block
// reductions and the lane-split CSR row product.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define KB 256

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// block sums of (a, b) in a fixed order; valid in thread 0
__device__ __forceinline__ void block_sum2(double &a, double &b)
{
    __shared__ double sa[KB / 64], sb[KB / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
#pragma unroll
        for (int k = 0; k < KB / 64; ++k) {
            x += sa[k];
            y += sb[k];
        }
        a = x;
        b = y;
    }
    __syncthreads();
}

// Nonzeros staged per block by csr_rows_lds (32 KB of LDS in FP64).
#define SPMV_CAP 4096
#ifndef SPMV_UNROLL
#define SPMV_UNROLL 8
#endif

// Rows r0 .. r1-1 (r1 - r0 <= blockDim.x) of y = A x.  The block streams its
// contiguous nonzero range rowptr[r0] .. rowptr[r1] through LDS as products
// v[k] * x[col[k]] (coalesced loads of values and columns), then thread t sums
// row r0 + t in column order.  Every thread of the block must call it with the
// same r0, r1; the result is valid for threads with r0 + t < r1.
// CAP is the number of entries lds holds.
template <class T, class V, class X, int CAP = SPMV_CAP>
__device__ __forceinline__ T csr_rows_lds(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                          const V *__restrict__ v, const X *__restrict__ x, int64_t r0, int64_t r1,
                                          T *__restrict__ lds)
{
    const int t = threadIdx.x;
    const int64_t r = r0 + t;
    const int32_t base = rp[r0], end = rp[r1];
    int32_t my0 = 0, my1 = 0;
    if (r < r1) {
        my0 = rp[r];
        my1 = rp[r + 1];
    }
    T s = (T)0;
    const int bd = blockDim.x;
    for (int32_t c0 = base; c0 < end; c0 += CAP) {
        const int32_t c1 = c0 + CAP < end ? c0 + CAP : end;
        __syncthreads();
        // SPMV_UNROLL products per step: all column and value loads, then all
        // x gathers, are issued before the first is used.  Loads past the
        // chunk's end are masked off (whole waves skip them).
        constexpr int U = SPMV_UNROLL;
        for (int32_t k = c0 + t; k < c1; k += U * bd) {
            int32_t c[U];
            V w[U];
            X xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = 0;
                w[u] = (V)0;
                if (k + u * bd < c1) {
                    c[u] = ci[k + u * bd];
                    w[u] = v[k + u * bd];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xv[u] = (X)0;
                if (k + u * bd < c1) xv[u] = x[c[u]];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + u * bd < c1) lds[k + u * bd - c0] = (T)w[u] * (T)xv[u];
        }
        __syncthreads();
        const int32_t a = my0 > c0 ? my0 : c0, b = my1 < c1 ? my1 : c1;
        for (int32_t k = a; k < b; ++k) s += lds[k - c0];
    }
    return s;
}

// FP64 staging at half the entries (16 KB of LDS): twice as many workgroups
// share a CU, which measured faster for every FP64 row product here.  The
// sum order does not depend on the staging size.
#define SPMV_CAP64 (SPMV_CAP / 2)
template <class V, class X>
__device__ __forceinline__ double csr_rows_lds64(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                 const V *__restrict__ v, const X *__restrict__ x, int64_t r0,
                                                 int64_t r1, double *__restrict__ lds)
{
    return csr_rows_lds<double, V, X, SPMV_CAP64>(rp, ci, v, x, r0, r1, lds);
}
