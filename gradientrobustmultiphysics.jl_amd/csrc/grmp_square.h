// Closed-form numbering of the generated unit-square triangulation: N x N
// squares, node (i, j) = j (N + 1) + i, square (i, j) split along its
// (0,0)-(1,1) diagonal into cell 2 (j N + i) (lower: (i,j), (i+1,j),
// (i+1,j+1)) and cell 2 (j N + i) + 1 (upper: (i,j), (i+1,j+1), (i,j+1)).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

// First-appearance face numbering of the unit-square triangulation (cells in
// order j, i, {lower, upper}; local faces [1 2; 2 3; 3 1]) in closed form:
// row 0 introduces 4N+1 faces, every later row 3N+1.
__device__ __host__ inline int64_t sq_base(int64_t N, int64_t j) { return j == 0 ? 0 : 4 * N + 1 + (j - 1) * (3 * N + 1); }
__device__ __host__ inline int64_t sq_H(int64_t N, int64_t i, int64_t jj)  // horizontal face at height jj
{
    if (jj == 0) return i == 0 ? 0 : 5 + 4 * (i - 1);
    int64_t r = jj - 1;
    if (r == 0) return i == 0 ? 3 : 5 + 4 * (i - 1) + 3;
    return sq_base(N, r) + (i == 0 ? 2 : 4 + 3 * (i - 1) + 2);
}
__device__ __host__ inline int64_t sq_V(int64_t N, int64_t ii, int64_t j)  // vertical face at x = ii in row j
{
    if (ii == 0) return j == 0 ? 4 : sq_base(N, j) + 3;
    if (j == 0) return ii == 1 ? 1 : 5 + 4 * (ii - 2) + 1;
    return ii == 1 ? sq_base(N, j) : sq_base(N, j) + 4 + 3 * (ii - 2);
}
__device__ __host__ inline int64_t sq_D(int64_t N, int64_t i, int64_t j)
{
    if (j == 0) return i == 0 ? 2 : 5 + 4 * (i - 1) + 2;
    return i == 0 ? sq_base(N, j) + 1 : sq_base(N, j) + 4 + 3 * (i - 1) + 1;
}

// CellFaces of cell c of the generated unit square (lower triangle of square
// (i,j) first, then the upper one)
__device__ __forceinline__ void sq_cellfaces(int64_t N, int64_t c, int64_t f[3])
{
    const int64_t sq = c >> 1, j = sq / N, i = sq % N;
    if (!(c & 1)) {
        f[0] = sq_H(N, i, j);
        f[1] = sq_V(N, i + 1, j);
        f[2] = sq_D(N, i, j);
    } else {
        f[0] = sq_D(N, i, j);
        f[1] = sq_H(N, i, j + 1);
        f[2] = sq_V(N, i, j);
    }
}
