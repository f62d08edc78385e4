// Element assembly of libgrmp_hip.so: sparsity pattern, BilinearForm,
// LinearForm and ItemIntegrator kernels for gfx950.
//
// Compiled with -ffp-contract=off (and the pragma below): every product and
// sum of the element math is rounded separately and in the order the
// reference's Julia loops use, so local values are bit-identical to the
// reference's on the same geometry.
//
// Design.  The reference loops over cells and scatters each local matrix
// into an ExtendableSparseMatrix (src/assemblypatterns/bilinearform.jl:226-377,
// _addnz src/fematrix.jl:54-65).  Here one thread owns one matrix ROW: it
// walks the cells that contain its dof (the space's incidence list, sorted by
// cell), recomputes that cell's local row, and adds each entry into an LDS
// copy of its row at a precomputed position (the 8-bit slot map).  Every
// entry is therefore summed in cell order -- the reference's order -- with no
// atomics, and each row is written to HBM once, coalesced across the block.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "grmp_internal.h"

#pragma clang fp contract(off)

#define RECO_MAXQ 25  // points of the largest triangle rule a kernel takes (Stroud order 9)
static int tri_rule(int order, double (*xq)[2], double *wq);

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#define MAXFORMS 4
// forms of one block row, kept in device memory (uniform loads in kernels)
struct FormPack {
    const DevForm *f;
    int n;
};

// ------------------------------------------------------- reference tables

// Triangle2D quadrature (src/quadrature.jl:173-195): NQ = 1 midpoint,
// NQ = 3 face midpoints [(1/2,1/2),(0,1/2),(1/2,0)], weights summing to 1.
template <int NQ>
__device__ __forceinline__ void qtri(int q, double &x0, double &x1, double &w)
{
    if (NQ == 1) {
        x0 = 1.0 / 3.0; x1 = 1.0 / 3.0; w = 1.0;
    } else {
        x0 = (q == 1) ? 0.0 : 0.5;
        x1 = (q == 2) ? 0.0 : 0.5;
        w = 1.0 / 3.0;
    }
}

// P1 triangle values (src/fedefs/h1_p1.jl:64-75)
__device__ __forceinline__ double p1tri_val(int a, double x0, double x1)
{
    double l = 1.0;
    l -= x0;
    l -= x1;
    return a == 0 ? l : (a == 1 ? x0 : x1);
}
__device__ __forceinline__ double p1tri_der(int a, int j)
{
    if (a == 0) return -1.0;
    return (a == j + 1) ? 1.0 : 0.0;
}
// P2 triangle values / exact derivatives (src/fedefs/h1_p2.jl:134-145)
__device__ __forceinline__ double p2tri_val(int a, double x0, double x1)
{
    double l = 1.0 - x0 - x1;
    switch (a) {
    case 0: return 2.0 * l * (l - 0.5);
    case 1: return 2.0 * x0 * (x0 - 0.5);
    case 2: return 2.0 * x1 * (x1 - 0.5);
    case 3: return 4.0 * l * x0;
    case 4: return 4.0 * x0 * x1;
    default: return 4.0 * x1 * l;
    }
}
__device__ __forceinline__ double p2tri_der(int a, int j, double x0, double x1)
{
    double l = 1.0 - x0 - x1;
    switch (a) {
    case 0: return -(4.0 * l - 1.0);
    case 1: return j == 0 ? 4.0 * x0 - 1.0 : 0.0;
    case 2: return j == 1 ? 4.0 * x1 - 1.0 : 0.0;
    case 3: return j == 0 ? 4.0 * (l - x0) : -4.0 * x0;
    case 4: return j == 0 ? 4.0 * x1 : 4.0 * x0;
    default: return j == 0 ? -4.0 * x1 : 4.0 * (l - x1);
    }
}

// ------------------------------------------------------------ affine maps
// Triangle: A = [x2-x1, x3-x1], det = A11 A22 - A12 A21,
// M = A^{-T} = [A22 -A21; -A12 A11]/det, vol = |det|/2
// (ExtendableGrids update_trafo!/mapderiv!, called at src/feevaluator.jl:371-380).
__device__ __forceinline__ void trafo_xy(double x0, double y0, double x1, double y1, double x2, double y2,
                                         double M[2][2], double &vol)
{
    double A11 = x1 - x0, A12 = x2 - x0, A21 = y1 - y0, A22 = y2 - y0;
    double det = A11 * A22 - A12 * A21;
    M[0][0] = A22 / det;
    M[0][1] = -A21 / det;
    M[1][0] = -A12 / det;
    M[1][1] = A11 / det;
    vol = fabs(det) * 0.5;
}

__device__ __forceinline__ void trafo_tri(const DevForm &F, int64_t e, double M[2][2], double &vol)
{
    const int64_t nc = F.ncells, nn = F.nnodes;
    int32_t v0 = F.cells[e], v1 = F.cells[nc + e], v2 = F.cells[2 * nc + e];
    trafo_xy(F.coords[v0], F.coords[nn + v0], F.coords[v1], F.coords[nn + v1], F.coords[v2], F.coords[nn + v2], M, vol);
}

__device__ __forceinline__ void trafo_x(const double X[4][3], double M[3][3], double &vol)
{
    double A[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) A[r][k] = X[k + 1][r] - X[0][r];
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    double c10 = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    double c11 = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    double c12 = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    double c20 = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    double c21 = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    double c22 = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    M[0][0] = c00 / det; M[0][1] = c01 / det; M[0][2] = c02 / det;
    M[1][0] = c10 / det; M[1][1] = c11 / det; M[1][2] = c12 / det;
    M[2][0] = c20 / det; M[2][1] = c21 / det; M[2][2] = c22 / det;
    vol = fabs(det) / 6.0;
}

__device__ __forceinline__ void trafo_tet(const DevForm &F, int64_t e, double M[3][3], double &vol)
{
    const int64_t nc = F.ncells, nn = F.nnodes;
    double X[4][3];
    for (int v = 0; v < 4; ++v) {
        int32_t n = F.cells[v * nc + e];
        X[v][0] = F.coords[n];
        X[v][1] = F.coords[nn + n];
        X[v][2] = F.coords[2 * nn + n];
    }
    trafo_x(X, M, vol);
}

// Corner coordinates and cell coefficient of one cell, gathered in two
// dependent steps (node ids, then coordinates and coefficient) so that a
// thread can keep the gathers of several cells in flight at once.  The
// row functions below take them and compute exactly what they compute from
// (F, e).
struct TriGeo {
    double x[3], y[3], kap;
};
struct TetGeo {
    double X[4][3], kap;
};

template <int NV>
__device__ __forceinline__ void cell_nodes(const DevForm &F, int64_t e, int32_t *nv)
{
#pragma unroll
    for (int v = 0; v < NV; ++v) nv[v] = F.cells[(int64_t)v * F.ncells + e];
}

__device__ __forceinline__ void geo_load(const DevForm &F, int64_t e, const int32_t *nv, TriGeo &g)
{
    const int64_t nn = F.nnodes;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        g.x[v] = F.coords[nv[v]];
        g.y[v] = F.coords[nn + nv[v]];
    }
    g.kap = F.coef_cell ? F.kappa[e] : 1.0;
}

__device__ __forceinline__ void geo_load(const DevForm &F, int64_t e, const int32_t *nv, TetGeo &g)
{
    const int64_t nn = F.nnodes;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        g.X[v][0] = F.coords[nv[v]];
        g.X[v][1] = F.coords[nn + nv[v]];
        g.X[v][2] = F.coords[2 * nn + nv[v]];
    }
    g.kap = F.coef_cell ? F.kappa[e] : 1.0;
}

// ----------------------------------------------------------- local rows
// Each row function writes v[b] = localmatrix(a,b) * itemfactor for the
// local row a of cell e: the value the reference passes to _addnz for the
// entry (dof_a, dof_b) of this cell.  For the symmetric pattern type the
// value of (a,b) is the upper-triangle entry (min, max) with the action
// applied to the min dof (bilinearform.jl:196-210, 329-346).  Loops over b
// are unrolled at compile time; the runtime index a only selects values.

// gradient of local scalar dof d (runtime) at a quadrature point:
// s_k = (0 + M[k][0] D[d][0]) + M[k][1] D[d][1]   (feevaluator_h1.jl:68-72)
template <int FE>
__device__ __forceinline__ void grad_tri(const double M[2][2], int d, double x0, double x1, double g[2])
{
    double d0 = (FE == 1) ? p1tri_der(d, 0) : p2tri_der(d, 0, x0, x1);
    double d1 = (FE == 1) ? p1tri_der(d, 1) : p2tri_der(d, 1, x0, x1);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double s = 0.0;
        s += M[k][0] * d0;
        s += M[k][1] * d1;
        g[k] = s;
    }
}

// (kappa grad u, grad v) with P1/P2 on triangles; NQ quadrature points.
template <int FE, int NQ>
__device__ __forceinline__ void row_laplace_tri_g(const DevForm &F, const TriGeo &G, int a, double *v)
{
    constexpr int ND = (FE == 1) ? 3 : 6;
    double M[2][2], vol;
    trafo_xy(G.x[0], G.y[0], G.x[1], G.y[1], G.x[2], G.y[2], M, vol);
    const double kap = G.kap;
    const double itemfactor = vol * F.factor * 1.0;
    double ga[NQ][2];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x0, x1, w;
        qtri<NQ>(q, x0, x1, w);
        grad_tri<FE>(M, a, x0, x1, ga[q]);
    }
#pragma unroll
    for (int b = 0; b < ND; ++b) {
        double L = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x0, x1, w;
            qtri<NQ>(q, x0, x1, w);
            double gb[2];
            grad_tri<FE>(M, b, x0, x1, gb);
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double lo_k = (a <= b) ? ga[q][k] : gb[k];
                double hi_k = (a <= b) ? gb[k] : ga[q][k];
                double act = F.coef_cell ? kap * lo_k : lo_k;
                t += act * hi_k;
            }
            L += w * t;
        }
        v[b] = L * itemfactor;
    }
}

// (kappa grad u, grad v) with P1 on tetrahedra, midpoint rule.
__device__ __forceinline__ void grad_tet(const double M[3][3], int d, double g[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double dj = (d == 0) ? -1.0 : ((d == j + 1) ? 1.0 : 0.0);
            s += M[k][j] * dj;
        }
        g[k] = s;
    }
}

// grad_tet for d = 0..3 at once, with the same bits when every entry of M is
// finite: its sum starts at +0, so the running sum is never -0 and the
// products M[k][j] * 0 (zeros of either sign) leave it unchanged; what
// remains is the sum from +0 of the terms with coefficient -1 or 1
__device__ __forceinline__ void grad_tet_all(const double M[3][3], double g[4][3])
{
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) fin &= isfinite(M[k][j]);
    if (!fin) {
#pragma unroll
        for (int d = 0; d < 4; ++d) grad_tet(M, d, g[d]);
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        g[0][k] = ((0.0 - M[k][0]) - M[k][1]) - M[k][2];
        g[1][k] = 0.0 + M[k][0];
        g[2][k] = 0.0 + M[k][1];
        g[3][k] = 0.0 + M[k][2];
    }
}

__device__ __forceinline__ void row_laplace_tet_g(const DevForm &F, const TetGeo &G, int a, double *v)
{
    double M[3][3], vol;
    trafo_x(G.X, M, vol);
    const double kap = G.kap;
    const double itemfactor = vol * F.factor * 1.0;
    double g4[4][3];
    grad_tet_all(M, g4);
    double ga[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) ga[k] = a == 0 ? g4[0][k] : a == 1 ? g4[1][k] : a == 2 ? g4[2][k] : g4[3][k];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const double *gb = g4[b];
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double lo_k = (a <= b) ? ga[k] : gb[k];
            double hi_k = (a <= b) ? gb[k] : ga[k];
            double act = F.coef_cell ? kap * lo_k : lo_k;
            t += act * hi_k;
        }
        double L = 0.0;
        L += 1.0 * t;
        v[b] = L * itemfactor;
    }
}

// (C eps(u), eps(v)) on P1{2} triangles (HookStiffnessOperator2D,
// src/pdeoperators.jl:256-276), pattern type APT_BilinearForm: the action
// acts on the row dof.  The per-cell coefficient scales both Lame
// parameters.  SymmetricGradient{1} Voigt values (feevaluator_h1.jl:97-116,
// compression [1,3,3,2]): a dof of component 0 gives [gx, 0, gy], of
// component 1 gives [0, gy, gx].
template <int NQ>
__device__ __forceinline__ void row_hooke_tri_g(const DevForm &F, const TriGeo &G, int a, double *v)
{
    double M[2][2], vol;
    trafo_xy(G.x[0], G.y[0], G.x[1], G.y[1], G.x[2], G.y[2], M, vol);
    double kap = G.kap;
    double lam = kap * F.lam, mu = kap * F.mu;
    const int ca = a / 3, sa = a - 3 * ca;
    double ga[2];
    grad_tri<1>(M, sa, 0.0, 0.0, ga);
    double in0 = ca == 0 ? ga[0] : 0.0, in1 = ca == 0 ? 0.0 : ga[1], in2 = ca == 0 ? ga[1] : ga[0];
    double r0 = (lam + 2.0 * mu) * in0 + lam * in1;
    double r1 = (lam + 2.0 * mu) * in1 + lam * in0;
    double r2 = mu * in2;
    const double itemfactor = vol * F.factor * 1.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const int cb = b / 3, sb = b % 3;
        double gb[2];
        grad_tri<1>(M, sb, 0.0, 0.0, gb);
        double c0 = cb == 0 ? gb[0] : 0.0, c1 = cb == 0 ? 0.0 : gb[1], c2 = cb == 0 ? gb[1] : gb[0];
        double t = 0.0;
        t += r0 * c0;
        t += r1 * c1;
        t += r2 * c2;
        double L = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x0, x1, w;
            qtri<NQ>(q, x0, x1, w);
            L += w * t;
        }
        v[b] = L * itemfactor;
    }
}

// (kappa u, v) with P1/P2 on triangles (ReactionOperator, symmetric).
template <int FE, int NQ>
__device__ __forceinline__ void row_mass_tri_g(const DevForm &F, const TriGeo &G, int a, double *v)
{
    constexpr int ND = (FE == 1) ? 3 : 6;
    double M[2][2], vol;
    trafo_xy(G.x[0], G.y[0], G.x[1], G.y[1], G.x[2], G.y[2], M, vol);
    const double kap = G.kap;
    const double itemfactor = vol * F.factor * 1.0;
#pragma unroll
    for (int b = 0; b < ND; ++b) {
        double L = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x0, x1, w;
            qtri<NQ>(q, x0, x1, w);
            double pa = (FE == 1) ? p1tri_val(a, x0, x1) : p2tri_val(a, x0, x1);
            double pb = (FE == 1) ? p1tri_val(b, x0, x1) : p2tri_val(b, x0, x1);
            double plo = (a <= b) ? pa : pb, phi = (a <= b) ? pb : pa;
            double act = F.coef_cell ? kap * plo : plo;
            double t = 0.0;
            t += act * phi;
            L += w * t;
        }
        v[b] = L * itemfactor;
    }
}

// ------------------------------------------------------- Bernardi-Raugel
// H1BR{2} (src/fedefs/h1v_br.jl:117-130,150-162): P1 node functions of both
// components, then face bubbles 6 l_a l_b times the global face normal.
// FaceNormals are restated from the face's global orientation (the first
// cell that met it, CellFaceSigns = +1): n = (t_y, -t_x)/|t|.
__device__ __forceinline__ void br_normals(const DevForm &F, int64_t e, double n[3][2])
{
    const int64_t nc = F.ncells, nn = F.nnodes;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int32_t a = F.cells[k * nc + e], b = F.cells[((k + 1) % 3) * nc + e];
        const bool own = F.fsign[k * nc + e] > 0;
        const int32_t ag = own ? a : b, bg = own ? b : a;
        double tx = F.coords[bg] - F.coords[ag], ty = F.coords[nn + bg] - F.coords[nn + ag];
        double len = sqrt(tx * tx + ty * ty);
        n[k][0] = ty / len;
        n[k][1] = -tx / len;
    }
}

// reference gradient of BR scalar function sf (0..2 nodes, 3..5 bubbles
// 6 x1 l1, 6 x1 x2, 6 l1 x2 with l1 = 1 - x1 - x2)
__device__ __forceinline__ void br_ref_der(int sf, double x0, double x1, double &d0, double &d1)
{
    double l = 1.0;
    l -= x0;
    l -= x1;
    if (sf < 3) {
        d0 = p1tri_der(sf, 0);
        d1 = p1tri_der(sf, 1);
    } else if (sf == 3) {
        d0 = 6.0 * (l - x0);
        d1 = -6.0 * x0;
    } else if (sf == 4) {
        d0 = 6.0 * x1;
        d1 = 6.0 * x0;
    } else {
        d0 = -6.0 * x1;
        d1 = 6.0 * (l - x1);
    }
}

// Gradient with coefficients (feevaluator_h1.jl:77-94) of BR dof `dof`:
// g[k + 2c] = (sum_j M[k][j] D_c[j]) * coefficient[c]
__device__ __forceinline__ void br_grad(const double M[2][2], const double n[3][2], int dof, double x0, double x1,
                                        double g[4])
{
    const int sf = dof < 6 ? dof % 3 : dof - 3;
    double d0, d1;
    br_ref_der(sf, x0, x1, d0, d1);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const bool active = dof < 6 ? (dof / 3 == c) : true;
        const double coef = dof < 6 ? 1.0 : (dof == 6 ? n[0][c] : (dof == 7 ? n[1][c] : n[2][c]));
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            double s = 0.0;
            s += M[k][0] * (active ? d0 : 0.0);
            s += M[k][1] * (active ? d1 : 0.0);
            s *= coef;
            g[k + 2 * c] = s;
        }
    }
}

// Divergence with coefficients (feevaluator_h1.jl:129-145)
__device__ __forceinline__ double br_div(const double M[2][2], const double n[3][2], int dof, double x0, double x1)
{
    const int sf = dof < 6 ? dof % 3 : dof - 3;
    double d0, d1;
    br_ref_der(sf, x0, x1, d0, d1);
    double dv = 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const bool active = dof < 6 ? (dof / 3 == k) : true;
        const double coef = dof < 6 ? 1.0 : (dof == 6 ? n[0][k] : (dof == 7 ? n[1][k] : n[2][k]));
        dv += M[k][0] * (active ? d0 : 0.0) * coef;
        dv += M[k][1] * (active ? d1 : 0.0) * coef;
    }
    return dv;
}

// LaplaceOperator(nu) on H1BR{2}: symmetric, NoAction, face-midpoint rule
template <int NQ>
__device__ __forceinline__ void row_laplace_br(const DevForm &F, int64_t e, int a, double *v)
{
    double M[2][2], vol, n[3][2];
    trafo_tri(F, e, M, vol);
    br_normals(F, e, n);
    const double itemfactor = vol * F.factor * 1.0;
    double ga[NQ][4];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x0, x1, w;
        qtri<NQ>(q, x0, x1, w);
        br_grad(M, n, a, x0, x1, ga[q]);
    }
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        double L = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x0, x1, w;
            qtri<NQ>(q, x0, x1, w);
            double gb[4];
            br_grad(M, n, b, x0, x1, gb);
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) t += (a <= b ? ga[q][k] * gb[k] : gb[k] * ga[q][k]);
            L += w * t;
        }
        v[b] = L * itemfactor;
    }
}

// LagrangeMultiplier(Divergence) block [BR rows, P0 column]: midpoint rule,
// P0 basis 1, itemfactor = vol * factor (factor -1)
__device__ __forceinline__ void row_div_brp0(const DevForm &F, int64_t e, int a, double *v)
{
    double M[2][2], vol, n[3][2];
    trafo_tri(F, e, M, vol);
    br_normals(F, e, n);
    double x0, x1, w;
    qtri<1>(0, x0, x1, w);
    double t = 0.0;
    t += br_div(M, n, a, x0, x1) * 1.0;
    double L = 0.0;
    L += w * t;
    v[0] = L * (vol * F.factor * 1.0);
}

// its transposed copy [P0 row, BR columns]:
// (L * itemfactor / factor * factor_t) * -1 (bilinearform.jl:358-363)
__device__ __forceinline__ void row_divt_p0br(const DevForm &F, int64_t e, int a, double *v)
{
    (void)a;
    double M[2][2], vol, n[3][2];
    trafo_tri(F, e, M, vol);
    br_normals(F, e, n);
    double x0, x1, w;
    qtri<1>(0, x0, x1, w);
    const double itemfactor = vol * F.factor * 1.0;
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        double t = 0.0;
        t += br_div(M, n, b, x0, x1) * 1.0;
        double L = 0.0;
        L += w * t;
        v[b] = (L * itemfactor / F.factor * F.factor_t) * -1.0;
    }
}

// ------------------------------------------------------------ Newton form
// NonlinearForm(Gradient, [Identity, Gradient], [1,1], kernel, [2,3]) with
// kernel(u, g) = alpha(u) g, alpha(u) = c0 + c1 u + c2 u^2
// (src/pdeoperators.jl:590-635): full_assemble! (src/assemblypatterns/
// nonlinearform.jl:44-246) evaluates the iterate's [u, grad u] at each
// point (eval_febe!, Identity then Gradient), the kernel's Jacobian
//   J = [alpha'(u) g_0, alpha, 0; alpha'(u) g_1, 0, alpha],
// and adds (J [phi_b, grad phi_b]) . grad phi_a w for ansatz b, test a;
// transposed_assembly puts it in row a (the test function), column b.
__device__ __forceinline__ void nl_point(const double *uc, int nd, int fe, const double M[2][2], double x0, double x1,
                                         double in[3])
{
    in[0] = in[1] = in[2] = 0.0;
    for (int d = 0; d < nd; ++d) in[0] += uc[d] * (fe == 1 ? p1tri_val(d, x0, x1) : p2tri_val(d, x0, x1)) * 1.0;
    for (int d = 0; d < nd; ++d) {
        double g[2];
        if (fe == 1) grad_tri<1>(M, d, x0, x1, g);
        else grad_tri<2>(M, d, x0, x1, g);
        in[1] += uc[d] * g[0] * 1.0;
        in[2] += uc[d] * g[1] * 1.0;
    }
}

// J and kernel value at input in = [u, g0, g1]
__device__ __forceinline__ void nl_jac(const double c[3], const double in[3], double J[2][3], double val[2])
{
    const double u = in[0];
    double alpha = c[0];
    alpha += c[1] * u;
    alpha += c[2] * (u * u);
    double dalpha = c[1];
    dalpha += 2.0 * c[2] * u;
    J[0][0] = dalpha * in[1];
    J[0][1] = alpha;
    J[0][2] = 0.0;
    J[1][0] = dalpha * in[2];
    J[1][1] = 0.0;
    J[1][2] = alpha;
    val[0] = alpha * in[1];
    val[1] = alpha * in[2];
}

template <int FE>
__device__ __forceinline__ void row_nldiff_tri(const DevForm &F, int64_t e, int a, double *v)
{
    constexpr int ND = (FE == 1) ? 3 : 6;
    if (!F.uold) {  // pattern: every local (test, ansatz) pair
#pragma unroll
        for (int b = 0; b < ND; ++b) v[b] = 1.0;
        return;
    }
    double M[2][2], vol;
    trafo_tri(F, e, M, vol);
    const int64_t nc = F.ncells;
    double uc[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) uc[d] = F.uold[F.rdofs[(int64_t)d * nc + e]];
    double L[ND];
#pragma unroll
    for (int b = 0; b < ND; ++b) L[b] = 0.0;
    for (int q = 0; q < F.nqt; ++q) {
        const double x0 = F.qtab[3 * q], x1 = F.qtab[3 * q + 1], w = F.qtab[3 * q + 2];
        double in[3], J[2][3], val[2];
        nl_point(uc, ND, FE, M, x0, x1, in);
        nl_jac(F.nlc, in, J, val);
        double ga[2];
        grad_tri<FE>(M, a, x0, x1, ga);
#pragma unroll
        for (int b = 0; b < ND; ++b) {
            double gb[2];
            grad_tri<FE>(M, b, x0, x1, gb);
            const double in2[3] = {(FE == 1 ? p1tri_val(b, x0, x1) : p2tri_val(b, x0, x1)), gb[0], gb[1]};
            double temp = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double ar = 0.0;
                ar += J[k][0] * in2[0];
                ar += J[k][1] * in2[1];
                ar += J[k][2] * in2[2];
                temp += ar * ga[k];
            }
            L[b] += temp * w;
        }
    }
    const double itemfactor = vol * F.factor * 1.0;
#pragma unroll
    for (int b = 0; b < ND; ++b) v[b] = L[b] * itemfactor;
}

// form-kind traits: number of column dofs and the row function
template <int CN_, void (*ROW)(const DevForm &, int64_t, int, double *)>
struct Kind {
    static constexpr int CN = CN_;
    static constexpr bool SPLIT = false;  // no separate gather step
    static constexpr int NV = 1;
    static constexpr int ROW_NB = 1;
    using Geo = int;
    __device__ static __forceinline__ void row(const DevForm &F, int64_t e, int a, double *v) { ROW(F, e, a, v); }
    __device__ static __forceinline__ void rowg(const DevForm &, const Geo &, int, double *) {}
};
// kinds whose rows depend on the cell only through its corners and
// coefficient: row() gathers them, rowg() computes from a gathered cell.
// RNB: incidences whose gathers the serial row walk keeps in flight together
// (measured fastest: 4 for tetrahedra's rows of about 24 cells, 3 for
// triangles' rows of 6, 2 for P2's rows of 1 to 6).
template <int CN_, class GEO, int NV_, void (*ROWG)(const DevForm &, const GEO &, int, double *), int RNB = 3>
struct KindG {
    static constexpr int CN = CN_;
    static constexpr bool SPLIT = true;
    static constexpr int NV = NV_;
    static constexpr int ROW_NB = RNB;
    using Geo = GEO;
    __device__ static __forceinline__ void rowg(const DevForm &F, const Geo &g, int a, double *v) { ROWG(F, g, a, v); }
    __device__ static __forceinline__ void row(const DevForm &F, int64_t e, int a, double *v)
    {
        int32_t nv[NV];
        cell_nodes<NV>(F, e, nv);
        Geo g;
        geo_load(F, e, nv, g);
        ROWG(F, g, a, v);
    }
};
using K_LAP_P1T1 = KindG<3, TriGeo, 3, row_laplace_tri_g<1, 1>>;
using K_LAP_P1T3 = KindG<3, TriGeo, 3, row_laplace_tri_g<1, 3>>;
using K_LAP_P2T3 = KindG<6, TriGeo, 3, row_laplace_tri_g<2, 3>, 2>;
using K_LAP_P1TET = KindG<4, TetGeo, 4, row_laplace_tet_g, 4>;
using K_HOOKE1 = KindG<6, TriGeo, 3, row_hooke_tri_g<1>>;
using K_HOOKE3 = KindG<6, TriGeo, 3, row_hooke_tri_g<3>>;
using K_MASS_P1T1 = KindG<3, TriGeo, 3, row_mass_tri_g<1, 1>>;
using K_MASS_P1T3 = KindG<3, TriGeo, 3, row_mass_tri_g<1, 3>>;
using K_LAP_BR = Kind<9, row_laplace_br<3>>;
using K_DIV_BRP0 = Kind<1, row_div_brp0>;
using K_DIVT_P0BR = Kind<9, row_divt_p0br>;
using K_NLDIFF_P1 = Kind<3, row_nldiff_tri<1>>;
using K_NLDIFF_P2 = Kind<6, row_nldiff_tri<2>>;

// runs the statement(s) with KT bound to the kind type of form F
#define GRMP_KCASE(CODE, TYPE, ...) \
    case CODE: {                    \
        using KT = TYPE;            \
        __VA_ARGS__;                \
    } break;
#define GRMP_DISPATCH(F, ...)                                                     \
    switch ((F).kind * 4 + (F).nq / 2) {                                          \
        GRMP_KCASE(FK_LAPLACE_P1TRI * 4 + 0, K_LAP_P1T1, __VA_ARGS__)             \
        GRMP_KCASE(FK_LAPLACE_P1TRI * 4 + 1, K_LAP_P1T3, __VA_ARGS__)             \
        GRMP_KCASE(FK_LAPLACE_P2TRI * 4 + 1, K_LAP_P2T3, __VA_ARGS__)             \
        GRMP_KCASE(FK_LAPLACE_P1TET * 4 + 0, K_LAP_P1TET, __VA_ARGS__)            \
        GRMP_KCASE(FK_HOOKE_P1VTRI * 4 + 0, K_HOOKE1, __VA_ARGS__)                \
        GRMP_KCASE(FK_HOOKE_P1VTRI * 4 + 1, K_HOOKE3, __VA_ARGS__)                \
        GRMP_KCASE(FK_MASS_P1TRI * 4 + 0, K_MASS_P1T1, __VA_ARGS__)               \
        GRMP_KCASE(FK_MASS_P1TRI * 4 + 1, K_MASS_P1T3, __VA_ARGS__)               \
        GRMP_KCASE(FK_LAPLACE_BR * 4 + 1, K_LAP_BR, __VA_ARGS__)                  \
        GRMP_KCASE(FK_DIV_BRP0 * 4 + 0, K_DIV_BRP0, __VA_ARGS__)                  \
        GRMP_KCASE(FK_DIVT_P0BR * 4 + 0, K_DIVT_P0BR, __VA_ARGS__)                \
        GRMP_KCASE(FK_NLDIFF_P1TRI * 4 + 0, K_NLDIFF_P1, __VA_ARGS__)             \
        GRMP_KCASE(FK_NLDIFF_P2TRI * 4 + 0, K_NLDIFF_P2, __VA_ARGS__)             \
    default: break;                                                               \
    }

// ---------------------------------------------------------------- pattern
// Per-thread sorted column set in LDS (layout set[k*B + tid]).
template <int B>
__device__ __forceinline__ bool set_insert(int32_t *set, int &n, int rmax, int32_t c)
{
    const int t = threadIdx.x;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        int32_t v = set[mid * B + t];
        if (v < c) lo = mid + 1;
        else hi = mid;
    }
    if (lo < n && set[lo * B + t] == c) return true;
    if (n >= rmax) return false;
    for (int k = n; k > lo; --k) set[k * B + t] = set[(k - 1) * B + t];
    set[lo * B + t] = c;
    ++n;
    return true;
}

template <int B>
__device__ __forceinline__ int set_find(const int32_t *set, int n, int32_t c)
{
    const int t = threadIdx.x;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (set[mid * B + t] < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// collects the columns with a nonzero local value of row i of form F
template <class K, int B>
__device__ __forceinline__ void pattern_collect(const DevForm &F, int64_t i, int32_t *set, int &n, int rmax, bool &ok)
{
    const int32_t b0 = F.inc_ptr[i], b1 = F.inc_ptr[i + 1];
    for (int32_t t = b0; t < b1; ++t) {
        const int32_t pk = F.inc[t];
        const int64_t e = pk / F.rnloc;
        const int a = pk - (int)e * F.rnloc;
        double v[K::CN];
        K::row(F, e, a, v);
#pragma unroll
        for (int b = 0; b < K::CN; ++b)
            if (v[b] != 0.0) ok &= set_insert<B>(set, n, rmax, (int32_t)(F.co + F.cdofs[(int64_t)b * F.ncells + e]));
    }
}

// writes the slot map of row i of form F: byte b of incidence t is the
// position of column cdofs[b] in the row, 255 for a zero local value
template <class K, int B>
__device__ __forceinline__ void pattern_slots(const DevForm &F, int64_t i, const int32_t *set, int n)
{
    const int32_t b0 = F.inc_ptr[i], b1 = F.inc_ptr[i + 1];
    constexpr int W = (K::CN + 3) / 4;
    for (int32_t t = b0; t < b1; ++t) {
        const int32_t pk = F.inc[t];
        const int64_t e = pk / F.rnloc;
        const int a = pk - (int)e * F.rnloc;
        double v[K::CN];
        K::row(F, e, a, v);
        uint32_t w[W];
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] = 0;
#pragma unroll
        for (int b = 0; b < K::CN; ++b) {
            uint32_t p = 255u;
            if (v[b] != 0.0) p = (uint32_t)set_find<B>(set, n, (int32_t)(F.co + F.cdofs[(int64_t)b * F.ncells + e]));
            w[b >> 2] |= p << (8 * (b & 3));
        }
#pragma unroll
        for (int q = 0; q < W; ++q) F.pos[(int64_t)t * W + q] = w[q];
    }
}

// FILL = false: count the row; FILL = true: write columns, diagonal
// position and the slot map of every (form, incidence, local column).
template <int B, int RMAX, bool FILL>
__global__ __launch_bounds__(B) void k_pattern(FormPack P, int64_t g0, int64_t g1, int64_t rblock,
                                               const uint8_t *__restrict__ fixed, int32_t *__restrict__ count,
                                               const int32_t *__restrict__ rowptr, int32_t *__restrict__ cols,
                                               uint8_t *__restrict__ diagpos, int *__restrict__ flag)
{
    __shared__ int32_t set[RMAX * B];
    const int64_t g = g0 + (int64_t)blockIdx.x * B + threadIdx.x;
    if (g >= g1) return;
    const int64_t i = g - rblock;
    int n = 0;
    bool ok = true;
    if (fixed[g]) ok &= set_insert<B>(set, n, RMAX, (int32_t)g);
    for (int f = 0; f < P.n; ++f) {
        const DevForm F = P.f[f];
        GRMP_DISPATCH(F, pattern_collect<KT, B>(F, i, set, n, RMAX, ok));
    }
    if (!ok) {
        atomicOr(flag, 1);
        return;
    }
    if (!FILL) {
        count[g] = n;
        atomicMax(flag + 1, n);
        return;
    }
    const int32_t r0 = rowptr[g];
    for (int k = 0; k < n; ++k) cols[r0 + k] = set[k * B + threadIdx.x];
    {
        int p = set_find<B>(set, n, (int32_t)g);
        diagpos[g] = (p < n && set[p * B + threadIdx.x] == (int32_t)g) ? (uint8_t)p : (uint8_t)255;
    }
    for (int f = 0; f < P.n; ++f) {
        const DevForm F = P.f[f];
        GRMP_DISPATCH(F, pattern_slots<KT, B>(F, i, set, n));
    }
}

// --------------------------------------------------------------- assembly
template <class K>
__device__ __forceinline__ void assemble_row(const DevForm &F, int64_t i, double *row)
{
    const int32_t b0 = F.inc_ptr[i], b1 = F.inc_ptr[i + 1];
    constexpr int W = (K::CN + 3) / 4;
    if constexpr (K::SPLIT) {
        // K::ROW_NB incidences at a time: their incidence and slot words,
        // node ids, then corners and coefficients are each loaded together
        // before any local row is computed; the row adds them in incidence
        // order as the loop below does.  Past the row's end a lane loads the
        // row's last incidence again and adds nothing.
        constexpr int NB = K::ROW_NB, NV = K::NV;
        for (int32_t t0 = b0; t0 < b1; t0 += NB) {
            int32_t pk[NB];
            uint32_t w[NB][W];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int32_t t = t0 + k < b1 ? t0 + k : b1 - 1;
                pk[k] = F.inc[t];
#pragma unroll
                for (int q = 0; q < W; ++q) w[k][q] = F.pos[(int64_t)t * W + q];
            }
            int64_t e[NB];
            int32_t nv[NB][NV];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                e[k] = pk[k] / F.rnloc;
                cell_nodes<NV>(F, e[k], nv[k]);
            }
            typename K::Geo G[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) geo_load(F, e[k], nv[k], G[k]);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (t0 + k < b1) {
                    double v[K::CN];
                    K::rowg(F, G[k], pk[k] - (int)e[k] * F.rnloc, v);
#pragma unroll
                    for (int b = 0; b < K::CN; ++b) {
                        const uint32_t p = (w[k][b >> 2] >> (8 * (b & 3))) & 255u;
                        if (p != 255u) row[p] += v[b];
                    }
                }
            }
        }
        return;
    }
    for (int32_t t = b0; t < b1; ++t) {
        const int32_t pk = F.inc[t];
        const int64_t e = pk / F.rnloc;
        const int a = pk - (int)e * F.rnloc;
        double v[K::CN];
        K::row(F, e, a, v);
        uint32_t w[W];
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] = F.pos[(int64_t)t * W + q];
#pragma unroll
        for (int b = 0; b < K::CN; ++b) {
            uint32_t p = (w[b >> 2] >> (8 * (b & 3))) & 255u;
            if (p != 255u) row[p] += v[b];
        }
    }
}

// The row block a workgroup owns.  With remap, consecutive row blocks go to
// one XCD: the hardware hands workgroup b to XCD b % 8, so workgroup b takes
// block i of XCD b % 8's contiguous share (the first nb % 8 XCDs get one
// block more); the cells and nodes that neighbouring rows share then meet in
// one L2 instead of eight.
__device__ __forceinline__ int64_t row_block_of(int remap)
{
    const int64_t b = blockIdx.x, nb = gridDim.x;
    if (!remap || nb < 16) return b;
    const int64_t x = b & 7, i = b >> 3, per = nb >> 3, rem = nb & 7;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// One thread per row; the block's rows are staged contiguously in LDS
// (dynamic shared memory of rowptr[end]-rowptr[start] doubles).  ``add``
// accumulates the forms of the row block into a row (FP64, reference order).
template <class AddRow>
__device__ __forceinline__ void assemble_block(AddRow add, int64_t blk, int64_t g0, int64_t g1, int64_t rblock,
                                               const int32_t *__restrict__ rowptr, const uint8_t *__restrict__ diagpos,
                                               const uint8_t *__restrict__ fixed, double penalty,
                                               double *__restrict__ val64, float *__restrict__ val32,
                                               double *__restrict__ dinv64, float *__restrict__ dinv32)
{
    extern __shared__ double buf[];
    const int B = blockDim.x;
    const int64_t rs = g0 + blk * B;
    const int64_t re = min(rs + B, g1);
    const int32_t base = rowptr[rs];
    const int32_t len = rowptr[re] - base;
    for (int32_t k = threadIdx.x; k < len; k += B) buf[k] = 0.0;
    __syncthreads();
    const int64_t g = rs + threadIdx.x;
    const bool mine = g < re;
    int32_t off = 0, rl = 0;
    bool fx = false;
    if (mine) {
        off = rowptr[g] - base;
        rl = rowptr[g + 1] - rowptr[g];
        fx = fixed[g] != 0;
        add(g - rblock, buf + off);
        const uint8_t dp = diagpos[g];
        if (fx && dp != 255) buf[off + dp] = penalty;
        double d = (dp != 255) ? buf[off + dp] : 0.0;
        double di = (fx || d == 0.0) ? 0.0 : 1.0 / d;
        if (dinv64) dinv64[g] = di;
        if (dinv32) dinv32[g] = (float)di;
    }
    __syncthreads();
    if (val64)
        for (int32_t k = threadIdx.x; k < len; k += B) val64[base + k] = buf[k];
    if (val32) {
        __syncthreads();
        if (mine && fx)
            for (int32_t k = 0; k < rl; ++k) buf[off + k] = 0.0;
        __syncthreads();
        for (int32_t k = threadIdx.x; k < len; k += B) val32[base + k] = (float)buf[k];
    }
}

// any number of forms per row block, kinds dispatched per form
__global__ void k_assemble(FormPack P, int64_t g0, int64_t g1, int64_t rblock, const int32_t *__restrict__ rowptr,
                           const uint8_t *__restrict__ diagpos, const uint8_t *__restrict__ fixed, double penalty,
                           double *__restrict__ val64, float *__restrict__ val32, double *__restrict__ dinv64,
                           float *__restrict__ dinv32)
{
    assemble_block(
        [&](int64_t i, double *row) {
            for (int f = 0; f < P.n; ++f) {
                const DevForm F = P.f[f];
                GRMP_DISPATCH(F, assemble_row<KT>(F, i, row));
            }
        },
        (int64_t)blockIdx.x, g0, g1, rblock, rowptr, diagpos, fixed, penalty, val64, val32, dinv64, dinv32);
}

// One form of a known kind, passed by value (kernel arguments stay scalar),
// with the cell work spread over the block.  The block owns rb <= ASM_NT rows
// and walks their contiguous incidence range in chunks:
//   1. every thread computes local rows of the chunk's incidences into LDS
//      (the gathers of all threads are in flight at once);
//   2. each row's thread adds its incidences of the chunk, in incidence
//      order, through the slot map into its LDS row.
// Per entry the summation order is the serial walk's (cell order), so the
// values are the same bits.
#define ASM_NT 256
#define ASM_STAGE 2048  // FP64 staging entries (16 KB)
template <class K>
__global__ __launch_bounds__(ASM_NT) void k_assemble1(DevForm F, int rb, int remap, int64_t g0, int64_t g1,
                                                      int64_t rblock,
                                                      const int32_t *__restrict__ rowptr,
                                                      const uint8_t *__restrict__ diagpos,
                                                      const uint8_t *__restrict__ fixed, double penalty,
                                                      double *__restrict__ val64, float *__restrict__ val32,
                                                      double *__restrict__ dinv64, float *__restrict__ dinv32)
{
    constexpr int CN = K::CN, W = (CN + 3) / 4, SCAP = ASM_STAGE / CN;
    extern __shared__ double shm[];
    double *stage = shm;             // SCAP * CN local values
    double *buf = shm + ASM_STAGE;   // the block's rows
    const int t = threadIdx.x;
    const int64_t rs = g0 + row_block_of(remap) * rb;
    const int64_t re = min(rs + rb, g1);
    const int32_t base = rowptr[rs];
    const int32_t len = rowptr[re] - base;
    for (int32_t k = t; k < len; k += ASM_NT) buf[k] = 0.0;
    const int64_t g = rs + t;
    const bool mine = t < rb && g < re;
    int32_t off = 0, rl = 0, my0 = 0, my1 = 0;
    if (mine) {
        off = rowptr[g] - base;
        rl = rowptr[g + 1] - rowptr[g];
        my0 = F.inc_ptr[g - rblock];
        my1 = F.inc_ptr[g - rblock + 1];
    }
    const int32_t i0 = F.inc_ptr[rs - rblock], i1 = F.inc_ptr[re - rblock];
    for (int32_t c0 = i0; c0 < i1; c0 += SCAP) {
        const int32_t c1 = min(c0 + SCAP, i1);
        __syncthreads();
        for (int32_t j = c0 + t; j < c1; j += ASM_NT) {
            const int32_t pk = F.inc[j];
            const int64_t e = pk / F.rnloc;
            const int a = pk - (int)e * F.rnloc;
            double v[CN];
            K::row(F, e, a, v);
#pragma unroll
            for (int b = 0; b < CN; ++b) stage[(j - c0) * CN + b] = v[b];
        }
        __syncthreads();
        if (mine) {
            const int32_t ja = max(my0, c0), jb = min(my1, c1);
            for (int32_t j = ja; j < jb; ++j) {
                uint32_t w[W];
#pragma unroll
                for (int q = 0; q < W; ++q) w[q] = F.pos[(int64_t)j * W + q];
#pragma unroll
                for (int b = 0; b < CN; ++b) {
                    const uint32_t p = (w[b >> 2] >> (8 * (b & 3))) & 255u;
                    if (p != 255u) buf[off + p] += stage[(j - c0) * CN + b];
                }
            }
        }
    }
    __syncthreads();
    bool fx = false;
    if (mine) {
        fx = fixed[g] != 0;
        const uint8_t dp = diagpos[g];
        if (fx && dp != 255) buf[off + dp] = penalty;
        const double d = (dp != 255) ? buf[off + dp] : 0.0;
        const double di = (fx || d == 0.0) ? 0.0 : 1.0 / d;
        if (dinv64) dinv64[g] = di;
        if (dinv32) dinv32[g] = (float)di;
    }
    __syncthreads();
    if (val64)
        for (int32_t k = t; k < len; k += ASM_NT) val64[base + k] = buf[k];
    if (val32) {
        __syncthreads();
        if (mine && fx)
            for (int32_t k = 0; k < rl; ++k) buf[off + k] = 0.0;
        __syncthreads();
        for (int32_t k = t; k < len; k += ASM_NT) val32[base + k] = (float)buf[k];
    }
}

// k_assemble1 with the cell gathers batched: in step 1 each thread takes NB
// incidences of the chunk (j = c0 + t + k ASM_NT) and issues all their
// loads level by level -- incidence words and slot words, node ids, corner
// coordinates and coefficients -- before computing any local row, so a
// thread waits for NB cells' memory once instead of NB times.  Step 1 also
// stages the slot words in LDS, so step 2 (each row's thread adding its
// incidences in incidence order) reads LDS only.  Summation order per entry
// is the serial walk's, so the values are the same bits.  Kinds with a
// separate gather step only.
#define ASMG_STAGE 3072  // FP64 staging entries (24 KB)
template <class K, int NB>
__global__ __launch_bounds__(ASM_NT) void k_assemble1g(DevForm F, int rb, int remap, int64_t g0, int64_t g1,
                                                       int64_t rblock, const int32_t *__restrict__ rowptr,
                                                       const uint8_t *__restrict__ diagpos,
                                                       const uint8_t *__restrict__ fixed, double penalty,
                                                       double *__restrict__ val64, float *__restrict__ val32,
                                                       double *__restrict__ dinv64, float *__restrict__ dinv32)
{
    constexpr int CN = K::CN, W = (CN + 3) / 4, NV = K::NV;
    constexpr int SCAP = (ASMG_STAGE / CN < NB * ASM_NT) ? ASMG_STAGE / CN : NB * ASM_NT;
    extern __shared__ double shm[];
    double *stage = shm;                                            // SCAP * CN local values
    uint32_t *wst = reinterpret_cast<uint32_t *>(shm + ASMG_STAGE);  // SCAP * W slot words
    double *buf = shm + ASMG_STAGE + (SCAP * W + 1) / 2;            // the block's rows
    const int t = threadIdx.x;
    const int64_t rs = g0 + row_block_of(remap) * rb;
    const int64_t re = min(rs + rb, g1);
    const int32_t base = rowptr[rs];
    const int32_t len = rowptr[re] - base;
    for (int32_t k = t; k < len; k += ASM_NT) buf[k] = 0.0;
    const int64_t g = rs + t;
    const bool mine = t < rb && g < re;
    int32_t off = 0, rl = 0, my0 = 0, my1 = 0;
    if (mine) {
        off = rowptr[g] - base;
        rl = rowptr[g + 1] - rowptr[g];
        my0 = F.inc_ptr[g - rblock];
        my1 = F.inc_ptr[g - rblock + 1];
    }
    const int32_t i0 = F.inc_ptr[rs - rblock], i1 = F.inc_ptr[re - rblock];
    for (int32_t c0 = i0; c0 < i1; c0 += SCAP) {
        const int32_t c1 = min(c0 + SCAP, i1);
        __syncthreads();
        for (int32_t jb = c0 + t; jb < c1; jb += NB * ASM_NT) {
            // past the chunk's end a lane repeats the chunk's last incidence and stores nothing
            int32_t pk[NB];
            uint32_t w[NB][W];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int32_t j = min(jb + k * ASM_NT, c1 - 1);
                pk[k] = F.inc[j];
#pragma unroll
                for (int q = 0; q < W; ++q) w[k][q] = F.pos[(int64_t)j * W + q];
            }
            int64_t e[NB];
            int32_t nv[NB][NV];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                e[k] = pk[k] / F.rnloc;
                cell_nodes<NV>(F, e[k], nv[k]);
            }
            typename K::Geo G[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) geo_load(F, e[k], nv[k], G[k]);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int32_t j = jb + k * ASM_NT;
                if (j < c1) {
                    const int a = pk[k] - (int)e[k] * F.rnloc;
                    double v[CN];
                    K::rowg(F, G[k], a, v);
#pragma unroll
                    for (int b = 0; b < CN; ++b) stage[(j - c0) * CN + b] = v[b];
#pragma unroll
                    for (int q = 0; q < W; ++q) wst[(j - c0) * W + q] = w[k][q];
                }
            }
        }
        __syncthreads();
        if (mine) {
            const int32_t ja = max(my0, c0), jb = min(my1, c1);
            for (int32_t j = ja; j < jb; ++j) {
#pragma unroll
                for (int b = 0; b < CN; ++b) {
                    const uint32_t p = (wst[(j - c0) * W + (b >> 2)] >> (8 * (b & 3))) & 255u;
                    if (p != 255u) buf[off + p] += stage[(j - c0) * CN + b];
                }
            }
        }
    }
    __syncthreads();
    bool fx = false;
    if (mine) {
        fx = fixed[g] != 0;
        const uint8_t dp = diagpos[g];
        if (fx && dp != 255) buf[off + dp] = penalty;
        const double d = (dp != 255) ? buf[off + dp] : 0.0;
        const double di = (fx || d == 0.0) ? 0.0 : 1.0 / d;
        if (dinv64) dinv64[g] = di;
        if (dinv32) dinv32[g] = (float)di;
    }
    __syncthreads();
    if (val64)
        for (int32_t k = t; k < len; k += ASM_NT) val64[base + k] = buf[k];
    if (val32) {
        __syncthreads();
        if (mine && fx)
            for (int32_t k = 0; k < rl; ++k) buf[off + k] = 0.0;
        __syncthreads();
        for (int32_t k = t; k < len; k += ASM_NT) val32[base + k] = (float)buf[k];
    }
}

// the LDS bytes of k_assemble1g<K, NB> with rb rows of at most mr entries
template <class K, int NB>
static size_t asmg_lds(int rb, int mr)
{
    constexpr int CN = K::CN, W = (CN + 3) / 4;
    constexpr int SCAP = (ASMG_STAGE / CN < NB * ASM_NT) ? ASMG_STAGE / CN : NB * ASM_NT;
    return (ASMG_STAGE + (SCAP * W + 1) / 2 + (size_t)rb * mr) * sizeof(double);
}

// One form of a known kind, one thread per row walking its cells serially:
// the better choice when rows have few incidences (triangles)
template <class K>
__global__ __launch_bounds__(256) void k_assemble1s(DevForm F, int remap, int64_t g0, int64_t g1, int64_t rblock,
                             const int32_t *__restrict__ rowptr, const uint8_t *__restrict__ diagpos,
                             const uint8_t *__restrict__ fixed, double penalty, double *__restrict__ val64,
                             float *__restrict__ val32, double *__restrict__ dinv64, float *__restrict__ dinv32)
{
    assemble_block([&](int64_t i, double *row) { assemble_row<K>(F, i, row); }, row_block_of(remap), g0, g1, rblock,
                   rowptr, diagpos, fixed, penalty, val64, val32, dinv64, dinv32);
}

int nnz_row_bucket(int max_row)
{
    return max_row <= 8 ? 0 : max_row <= 16 ? 1 : max_row <= 32 ? 2 : max_row <= 64 ? 3 : 4;
}

// the launch shape for a single-form row block of kind hf: the matrix's own,
// else the context's tuning table, else the built-in choice (variant and rows
// 0, remap -1 are filled in by launch_assemble)
static AsmTune asm_tune(const grmp_matrix *A, const DevForm &hf)
{
    if (A->tune_set) return A->tune;
    const int bucket = nnz_row_bucket(A->max_row);
    for (const TuneEntry &e : A->ctx->tune)
        if (e.kind == hf.kind && e.nq == hf.nq && e.bucket == bucket) return e.t;
    return AsmTune();
}

static int asm_block_size(int max_row);

// variant 3 for kinds with a separate gather step (NB cells in flight per
// thread: the stage holds NB * ASM_NT local rows); false for the others
template <class K>
static bool launch_assemble1g(const DevForm &F, int rb, int remap, int64_t g0, int64_t g1, const grmp_matrix *A,
                              int mr, double *val64, float *val32, double *dinv64, float *dinv32, hipStream_t st)
{
    if constexpr (K::SPLIT) {
        constexpr int NB = ASMG_STAGE / (K::CN * ASM_NT) > 0 ? ASMG_STAGE / (K::CN * ASM_NT) : 1;
        const size_t lds = asmg_lds<K, NB>(rb, mr);
        if (lds > 64 * 1024) return false;
        hipLaunchKernelGGL((k_assemble1g<K, NB>), dim3(nblk(g1 - g0, rb)), dim3(ASM_NT), lds, st, F, rb, remap, g0, g1,
                           g0, A->rowptr, A->diagpos, A->fixed, A->penalty, val64, val32, dinv64, dinv32);
        return true;
    } else {
        return false;
    }
}

// Launches the value assembly of row block [g0, g1): a per-kind kernel when
// the block has one form (hf = its host copy), the generic one otherwise.
// Built-in choice for one form: the serial row walk with batched gathers for
// kinds with a separate gather step; for the others the serial walk up to 12
// cell incidences per row, the staged kernel above; blocks in the hardware's
// order.
static void launch_assemble(const grmp_matrix *A, const grmp_space *sp, const DevForm *hf, FormPack P, int64_t g0,
                            int64_t g1, int B, size_t shm, double *val64, float *val32, double *dinv64, float *dinv32,
                            hipStream_t st)
{
    const double inc_per_row = (double)sp->mesh->ncells * sp->nloc / (double)std::max<int64_t>(sp->ndofs, 1);
    const int mr = A->max_row > 0 ? A->max_row : 1;
    if (P.n == 1) {
        const AsmTune T = asm_tune(A, *hf);
        bool split = false;
        {
            const DevForm F = *hf;
            GRMP_DISPATCH(F, split = KT::SPLIT);
        }
        const int variant = T.variant ? T.variant : ((split || inc_per_row <= 12.0) ? 1 : 3);
        const int remap = T.remap >= 0 ? T.remap : 0;
        if (variant == 1) {
            // rows per block: a multiple of 64 whose rows fit the 48 KB of LDS
            int Bs = asm_block_size(mr);
            if (T.rows >= 64 && T.rows <= Bs) Bs = T.rows & ~63;
            const DevForm F = *hf;
            const size_t shs = (size_t)Bs * mr * sizeof(double);
            GRMP_DISPATCH(F, hipLaunchKernelGGL((k_assemble1s<KT>), dim3(nblk(g1 - g0, Bs)), dim3(Bs), shs, st, F, remap, g0,
                                                g1, g0, A->rowptr, A->diagpos, A->fixed, A->penalty, val64, val32,
                                                dinv64, dinv32));
            return;
        }
        // rows per block: about 6 incidences per thread in a chunk, and the
        // block's rows (rb * max_row doubles) within 48 KB next to the 16 KB stage
        int rb = ASM_NT;
        while (rb > 32 && ((double)rb * inc_per_row > 6.0 * ASM_NT || (size_t)rb * mr * 8 > 48 * 1024)) rb >>= 1;
        if (T.rows >= 8 && T.rows <= ASM_NT && (size_t)T.rows * mr * 8 <= 48 * 1024) rb = T.rows;
        if (variant == 3 && (size_t)rb * mr * 8 <= 40 * 1024) {
            const DevForm F = *hf;
            bool done = false;
            GRMP_DISPATCH(F, done = launch_assemble1g<KT>(F, rb, remap, g0, g1, A, mr, val64, val32, dinv64, dinv32, st));
            if (done) return;
        }
        if ((size_t)rb * mr * 8 <= 48 * 1024) {
            const DevForm F = *hf;
            const size_t shm1 = (ASM_STAGE + (size_t)rb * mr) * sizeof(double);
            GRMP_DISPATCH(F, hipLaunchKernelGGL((k_assemble1<KT>), dim3(nblk(g1 - g0, rb)), dim3(ASM_NT), shm1, st, F, rb,
                                                remap, g0, g1, g0, A->rowptr, A->diagpos, A->fixed, A->penalty, val64,
                                                val32, dinv64, dinv32));
            return;
        }
    }
    hipLaunchKernelGGL(k_assemble, dim3(nblk(g1 - g0, B)), dim3(B), shm, st, P, g0, g1, g0, A->rowptr, A->diagpos,
                       A->fixed, A->penalty, val64, val32, dinv64, dinv32);
}

// ------------------------------------------------------------ matrix API

static int quad_nq_tri(int order) { return order <= 1 ? 1 : (order == 2 ? 3 : 0); }

int form_resolve(grmp_matrix *A, FormRec &r)
{
    grmp_ctx *ctx = A->ctx;
    const grmp_form &f = r.f;
    const grmp_space *rs = r.rs, *cs = r.cs;
    GRMP_CHECK(ctx, rs->mesh == cs->mesh, GRMP_ERR_ARG, "form: row and column spaces live on different meshes");
    const int dim = rs->mesh->dim;
    r.kind = 0;
    r.nq = 0;
    if (f.kind == GRMP_BLF_LAPLACE && rs == cs && rs->ncomp == 1) {
        if (dim == 2 && rs->fe == GRMP_FE_H1P1) { r.kind = FK_LAPLACE_P1TRI; r.nq = quad_nq_tri(f.quadorder); }
        if (dim == 2 && rs->fe == GRMP_FE_H1P2 && f.quadorder == 2) { r.kind = FK_LAPLACE_P2TRI; r.nq = 3; }
        if (dim == 3 && rs->fe == GRMP_FE_H1P1 && f.quadorder <= 1) { r.kind = FK_LAPLACE_P1TET; r.nq = 1; }
    } else if (f.kind == GRMP_BLF_HOOKE2D && rs == cs && rs->ncomp == 2 && dim == 2 && rs->fe == GRMP_FE_H1P1) {
        r.kind = FK_HOOKE_P1VTRI;
        r.nq = quad_nq_tri(f.quadorder);
    } else if (f.kind == GRMP_BLF_MASS && rs == cs && rs->ncomp == 1 && dim == 2 && rs->fe == GRMP_FE_H1P1) {
        r.kind = FK_MASS_P1TRI;
        r.nq = quad_nq_tri(f.quadorder);
    } else if (f.kind == GRMP_BLF_LAPLACE && rs == cs && rs->fe == GRMP_FE_H1BR && dim == 2 && f.quadorder == 2 &&
               f.coef_mode == GRMP_COEF_FACTOR) {
        r.kind = FK_LAPLACE_BR;
        r.nq = 3;
    } else if (f.kind == GRMP_BLF_DIVP0 && rs->fe == GRMP_FE_H1BR && cs->fe == GRMP_FE_L2P0 && dim == 2 &&
               f.quadorder <= 1 && f.coef_mode == GRMP_COEF_FACTOR) {
        r.kind = FK_DIV_BRP0;
        r.nq = 1;
    } else if (f.kind == GRMP_BLF_NLDIFF && rs == cs && rs->ncomp == 1 && dim == 2 &&
               (rs->fe == GRMP_FE_H1P1 || rs->fe == GRMP_FE_H1P2) && f.coef_mode == GRMP_COEF_FACTOR) {
        GRMP_CHECK(ctx, f.current && f.current_offset >= 0 && f.current_offset + rs->ndofs <= f.current->n,
                   GRMP_ERR_ARG, "Newton form: the iterate vector is missing or too short");
        double xq[RECO_MAXQ][2], wq[RECO_MAXQ];
        const int nqt = tri_rule(f.quadorder, xq, wq);
        GRMP_CHECK(ctx, nqt > 0, GRMP_ERR_UNSUPPORTED, "Newton form: quadrature order %d too high", f.quadorder);
        std::vector<double> tab(3 * nqt);
        for (int q = 0; q < nqt; ++q) {
            tab[3 * q] = xq[q][0];
            tab[3 * q + 1] = xq[q][1];
            tab[3 * q + 2] = wq[q];
        }
        GRMP_HIP(ctx, hipMalloc((void **)&r.qtab, sizeof(double) * tab.size()));
        GRMP_HIP(ctx, hipMemcpy(r.qtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
        r.nqt = nqt;
        r.kind = rs->fe == GRMP_FE_H1P1 ? FK_NLDIFF_P1TRI : FK_NLDIFF_P2TRI;
        r.nq = 1;  // dispatch key; the points are in the table
    }
    GRMP_CHECK(ctx, (r.kind != FK_LAPLACE_BR && r.kind != FK_DIV_BRP0) || rs->mesh->fsign, GRMP_ERR_UNSUPPORTED,
               "H1BR forms need the mesh's CellFaceSigns (generated meshes carry them)");
    GRMP_CHECK(ctx, r.kind != 0 && r.nq != 0, GRMP_ERR_UNSUPPORTED,
               "form kind %d (quadorder %d) on FE %d/%d with %d components in %dD is not built", f.kind,
               f.quadorder, rs->fe, cs->fe, rs->ncomp, dim);
    GRMP_CHECK(ctx, f.coef_mode == GRMP_COEF_FACTOR || (f.coef_mode == GRMP_COEF_CELL && f.cell_coef &&
                                                         f.cell_coef->n == rs->mesh->ncells),
               GRMP_ERR_ARG, "form: GRMP_COEF_CELL needs a cell_coef vector of length ncells");
    r.words = (cs->nloc + 3) / 4;
    return GRMP_OK;
}

void fill_devform(const grmp_matrix *A, const FormRec &r, DevForm &d)
{
    memset(&d, 0, sizeof(d));
    const grmp_mesh *m = r.rs->mesh;
    d.kind = r.kind;
    d.nq = r.nq;
    d.coef_cell = r.f.coef_mode == GRMP_COEF_CELL;
    d.rnloc = r.rs->nloc;
    d.cnloc = r.cs->nloc;
    d.words = r.words;
    d.factor = r.f.factor;
    d.factor_t = r.f.factor;
    d.fsign = m->fsign;
    d.lam = r.f.params[0];
    d.mu = r.f.params[1];
    d.kappa = d.coef_cell ? r.f.cell_coef->d : nullptr;
    d.coords = m->coords;
    d.cells = m->cells;
    d.nnodes = m->nnodes;
    d.ncells = m->ncells;
    d.rdofs = r.rs->celldofs;
    d.inc_ptr = r.rs->inc_ptr;
    d.inc = r.rs->inc;
    d.cdofs = r.cs->celldofs;
    d.ro = A->offs[r.f.row_block];
    d.co = A->offs[r.f.col_block];
    d.pos = r.pos;
    if (r.kind == FK_NLDIFF_P1TRI || r.kind == FK_NLDIFF_P2TRI) {
        d.uold = r.f.current->d + r.f.current_offset;
        d.qtab = r.qtab;
        d.nqt = r.nqt;
        for (int k = 0; k < 3; ++k) d.nlc[k] = r.f.params[k];
    }
}

extern "C" int grmp_matrix_create(grmp_ctx *ctx, int nblocks, const grmp_space *const *spaces, grmp_matrix **out)
{
    GRMP_CHECK(ctx, ctx && out && spaces && nblocks >= 1 && nblocks <= 8, GRMP_ERR_ARG, "grmp_matrix_create: bad arguments");
    grmp_matrix *A = new grmp_matrix();
    A->ctx = ctx;
    A->nblocks = nblocks;
    A->offs[0] = 0;
    for (int b = 0; b < nblocks; ++b) {
        if (!spaces[b]) { delete A; return grmp_fail(ctx, GRMP_ERR_ARG, "grmp_matrix_create: null space"); }
        A->spaces[b] = spaces[b];
        A->offs[b + 1] = A->offs[b] + spaces[b]->ndofs;
    }
    A->n = A->offs[nblocks];
    if (A->n >= (1ll << 31)) { delete A; return grmp_fail(ctx, GRMP_ERR_ARG, "grmp_matrix_create: too many dofs"); }
    *out = A;
    return GRMP_OK;
}

static void matrix_free_pattern(grmp_matrix *A)
{
    grmp_ctx *ctx = A->ctx;
    dfree(ctx, A->rowptr);
    dfree(ctx, A->cols);
    dfree(ctx, A->diagpos);
    dfree(ctx, A->val64);
    dfree(ctx, A->val32);
    dfree(ctx, A->dinv64);
    dfree(ctx, A->dinv32);
    for (auto &r : A->forms) dfree(ctx, r.pos);
    A->have_pattern = false;
    A->values = 0;
}

extern "C" int grmp_matrix_destroy(grmp_matrix *A)
{
    if (!A) return GRMP_OK;
    std::lock_guard<std::mutex> g(A->ctx->mu);
    matrix_free_pattern(A);
    for (auto &r : A->forms)
        if (r.qtab) (void)hipFree(r.qtab);
    dfree(A->ctx, A->fixed);
    delete A;
    return GRMP_OK;
}

extern "C" int grmp_matrix_add_form(grmp_matrix *A, const grmp_form *form)
{
    if (!A || !form) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, form->row_block >= 0 && form->row_block < A->nblocks && form->col_block >= 0 &&
                        form->col_block < A->nblocks,
               GRMP_ERR_ARG, "grmp_matrix_add_form: block index out of range");
    int same = 0;
    for (auto &r : A->forms) same += r.f.row_block == form->row_block;
    GRMP_CHECK(ctx, same < MAXFORMS, GRMP_ERR_UNSUPPORTED, "grmp_matrix_add_form: at most %d forms per block row", MAXFORMS);
    FormRec r;
    r.f = *form;
    r.rs = A->spaces[form->row_block];
    r.cs = A->spaces[form->col_block];
    GRMP_TRY(form_resolve(A, r));
    std::lock_guard<std::mutex> g(ctx->mu);
    matrix_free_pattern(A);
    A->forms.push_back(r);
    if (r.kind == FK_DIV_BRP0) {
        // LagrangeMultiplier: transposed copy into the mirrored block
        // (src/pdeoperators.jl:214-226, assemble! with At)
        FormRec t = r;
        t.f.row_block = r.f.col_block;
        t.f.col_block = r.f.row_block;
        t.rs = r.cs;
        t.cs = r.rs;
        t.kind = FK_DIVT_P0BR;
        t.nq = 1;
        t.pos = nullptr;
        t.words = (t.cs->nloc + 3) / 4;
        A->forms.push_back(t);
    }
    return GRMP_OK;
}

extern "C" int grmp_matrix_set_fixed(grmp_matrix *A, int64_t nfixed, const int64_t *fixed, double penalty)
{
    if (!A || nfixed < 0 || (nfixed && !fixed)) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    std::vector<uint8_t> mask(A->n, 0);
    for (int64_t k = 0; k < nfixed; ++k) {
        GRMP_CHECK(ctx, fixed[k] >= 1 && fixed[k] <= A->n, GRMP_ERR_ARG, "grmp_matrix_set_fixed: dof %lld out of range",
                   (long long)fixed[k]);
        mask[fixed[k] - 1] = 1;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    matrix_free_pattern(A);
    A->fixed_host.assign(fixed, fixed + nfixed);
    A->penalty = penalty;
    if (!A->fixed) GRMP_TRY(dalloc(ctx, &A->fixed, A->n));
    GRMP_HIP(ctx, hipMemcpyAsync(A->fixed, mask.data(), A->n, hipMemcpyHostToDevice, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GRMP_OK;
}

// uploads the DevForms of every block row (in registration order) and
// returns the device pack of block row rb
// the pattern pass uploads Newton forms without their iterate: every local
// pair then enters the pattern, whatever the iterate's values
static int upload_forms(grmp_matrix *A, bool pattern, const double *kappa_override = nullptr)
{
    grmp_ctx *ctx = A->ctx;
    std::vector<DevForm> h;
    A->form_first.assign(A->nblocks + 1, 0);
    for (int rb = 0; rb < A->nblocks; ++rb) {
        A->form_first[rb] = (int)h.size();
        for (auto &r : A->forms)
            if (r.f.row_block == rb) {
                DevForm d;
                fill_devform(A, r, d);
                if (kappa_override && d.coef_cell) d.kappa = kappa_override;
                if (pattern) d.uold = nullptr;
                h.push_back(d);
            }
    }
    A->form_first[A->nblocks] = (int)h.size();
    if (!A->dforms) GRMP_TRY(dalloc(ctx, &A->dforms, A->nblocks * MAXFORMS));
    GRMP_HIP(ctx, hipMemcpyAsync(A->dforms, h.data(), sizeof(DevForm) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    // the copy source is a host vector: finish it before returning
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GRMP_OK;
}

static FormPack row_block_pack(const grmp_matrix *A, int rb)
{
    FormPack P;
    P.f = A->dforms + A->form_first[rb];
    P.n = A->form_first[rb + 1] - A->form_first[rb];
    return P;
}

template <int B, int RMAX>
static int pattern_pass(grmp_matrix *A, bool fill, int32_t *count)
{
    grmp_ctx *ctx = A->ctx;
    for (int rb = 0; rb < A->nblocks; ++rb) {
        FormPack P = row_block_pack(A, rb);
        int64_t g0 = A->offs[rb], g1 = A->offs[rb + 1];
        if (g1 <= g0) continue;
        if (fill)
            hipLaunchKernelGGL((k_pattern<B, RMAX, true>), dim3(nblk(g1 - g0, B)), dim3(B), 0, ctx->stream, P, g0, g1,
                               g0, A->fixed, count, A->rowptr, A->cols, A->diagpos, ctx->dflag);
        else
            hipLaunchKernelGGL((k_pattern<B, RMAX, false>), dim3(nblk(g1 - g0, B)), dim3(B), 0, ctx->stream, P, g0, g1,
                               g0, A->fixed, count, A->rowptr, A->cols, A->diagpos, ctx->dflag);
        GRMP_HIP(ctx, hipGetLastError());
    }
    return GRMP_OK;
}

extern "C" int grmp_matrix_build_pattern(grmp_matrix *A)
{
    if (!A) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, !A->forms.empty(), GRMP_ERR_ARG, "grmp_matrix_build_pattern: no forms registered");
    std::lock_guard<std::mutex> g(ctx->mu);
    matrix_free_pattern(A);
    if (!A->fixed) {
        GRMP_TRY(dalloc(ctx, &A->fixed, A->n));
        GRMP_HIP(ctx, hipMemsetAsync(A->fixed, 0, A->n, ctx->stream));
    }
    for (auto &r : A->forms) GRMP_TRY(dalloc(ctx, &r.pos, (size_t)r.rs->mesh->ncells * r.rs->nloc * r.words));
    GRMP_TRY(upload_forms(A, true));
    int32_t *count = nullptr;
    GRMP_TRY(dalloc(ctx, &count, A->n));
    GRMP_TRY(dalloc(ctx, &A->rowptr, A->n + 1));
    GRMP_TRY(dalloc(ctx, &A->diagpos, A->n));
    int flags[2] = {0, 0};
    int big = 0;
    for (;; ++big) {
        GRMP_HIP(ctx, hipMemsetAsync(ctx->dflag, 0, 2 * sizeof(int), ctx->stream));
        if (big == 0) GRMP_TRY((pattern_pass<64, 64>(A, false, count)));
        else GRMP_TRY((pattern_pass<32, 255>(A, false, count)));
        GRMP_HIP(ctx, hipMemcpyAsync(flags, ctx->dflag, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (!flags[0]) break;
        if (big) {
            dfree(ctx, count);
            return grmp_fail(ctx, GRMP_ERR_UNSUPPORTED, "grmp_matrix_build_pattern: a row has more than 255 entries");
        }
    }
    A->max_row = flags[1];
    // the row pointers are int32: check the total in 64 bits before the scan
    int64_t total = 0;
    int rc64 = sum_i32_as_i64(ctx, count, A->n, &total);
    const int64_t limit = std::min<int64_t>(GRMP_MAX_NNZ, ctx->max_nnz);
    if (!rc64 && total > limit)
        rc64 = grmp_fail(ctx, GRMP_ERR_UNSUPPORTED, "grmp_matrix_build_pattern: %lld nonzeros exceed the limit %lld "
                         "(int32 CSR: %lld)", (long long)total, (long long)limit, (long long)GRMP_MAX_NNZ);
    if (rc64) {
        dfree(ctx, count);
        return rc64;
    }
    GRMP_TRY(exclusive_scan_i32(ctx, count, A->rowptr, A->n));
    int32_t nnz32 = 0;
    GRMP_HIP(ctx, hipMemcpyAsync(&nnz32, A->rowptr + A->n, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    A->nnz = nnz32;
    GRMP_TRY(dalloc(ctx, &A->cols, A->nnz));
    if (big == 0) GRMP_TRY((pattern_pass<64, 64>(A, true, count)));
    else GRMP_TRY((pattern_pass<32, 255>(A, true, count)));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, count);
    A->have_pattern = true;
    return GRMP_OK;
}

static int asm_block_size(int max_row)
{
    int B = 256;
    while (B > 32 && (size_t)B * max_row * sizeof(double) > 48 * 1024) B >>= 1;
    return B;
}

extern "C" int grmp_matrix_assemble(grmp_matrix *A, int values)
{
    if (!A) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    if (!A->have_pattern) GRMP_TRY(grmp_matrix_build_pattern(A));
    GRMP_CHECK(ctx, values & (GRMP_VALUES_FP64 | GRMP_VALUES_FP32), GRMP_ERR_ARG, "grmp_matrix_assemble: no value precision requested");
    std::lock_guard<std::mutex> g(ctx->mu);
    if ((values & GRMP_VALUES_FP64) && !A->val64) GRMP_TRY(dalloc(ctx, &A->val64, A->nnz));
    if ((values & GRMP_VALUES_FP32) && !A->val32) GRMP_TRY(dalloc(ctx, &A->val32, A->nnz));
    if (!A->dinv64) GRMP_TRY(dalloc(ctx, &A->dinv64, A->n));
    if (!A->dinv32) GRMP_TRY(dalloc(ctx, &A->dinv32, A->n));
    // Newton forms read the iterate: the device copies of the forms made for
    // the pattern (without it) are refreshed for the multi-form kernel
    bool newton = false;
    for (auto &r : A->forms) newton |= r.qtab != nullptr;
    if (newton) GRMP_TRY(upload_forms(A, false));
    const int B = asm_block_size(A->max_row);
    std::vector<DevForm> hf(A->forms.size() + 1);
    {
        int k = 0;
        for (int rb = 0; rb < A->nblocks; ++rb)
            for (auto &r : A->forms)
                if (r.f.row_block == rb) fill_devform(A, r, hf[k++]);
    }
    // the event pair lives until the function returns, whatever the path
    struct EventPair {
        hipEvent_t a = nullptr, b = nullptr;
        ~EventPair()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    GRMP_HIP(ctx, hipEventCreate(&ev.a));
    GRMP_HIP(ctx, hipEventCreate(&ev.b));
    GRMP_HIP(ctx, hipEventRecord(ev.a, ctx->stream));
    for (int rb = 0; rb < A->nblocks; ++rb) {
        FormPack P = row_block_pack(A, rb);
        int64_t g0 = A->offs[rb], g1 = A->offs[rb + 1];
        if (g1 <= g0) continue;
        size_t shm = (size_t)B * (A->max_row > 0 ? A->max_row : 1) * sizeof(double);
        launch_assemble(A, A->spaces[rb], &hf[A->form_first[rb]], P, g0, g1, B, shm,
                        (values & GRMP_VALUES_FP64) ? A->val64 : nullptr,
                        (values & GRMP_VALUES_FP32) ? A->val32 : nullptr, A->dinv64, A->dinv32, ctx->stream);
        GRMP_HIP(ctx, hipGetLastError());
    }
    GRMP_HIP(ctx, hipEventRecord(ev.b, ctx->stream));
    GRMP_HIP(ctx, hipEventSynchronize(ev.b));
    GRMP_HIP(ctx, hipEventElapsedTime(&A->last_asm_ms, ev.a, ev.b));
    A->values |= values;
    return GRMP_OK;
}

// Assembles the matrix values of A's pattern into the given buffers on
// stream st, with kappa replacing the cell coefficient of every
// GRMP_COEF_CELL form.  dforms: device array (>= forms.size()) owned by the
// caller; hforms: host staging that must stay valid until st passes it.
int assemble_values(const grmp_matrix *A, const double *kappa, DevForm *dforms, DevForm *hforms, double *val64,
                    float *val32, double *dinv64, float *dinv32, hipStream_t st)
{
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, A->have_pattern && A->rowptr && A->diagpos && A->fixed && dforms && hforms && dinv64 && dinv32 &&
                        (val64 || val32),
               GRMP_ERR_ARG, "assemble_values: pattern or output buffers missing");
    for (auto &r : A->forms)
        GRMP_CHECK(ctx, r.pos && (r.f.coef_mode != GRMP_COEF_CELL || kappa || r.f.cell_coef), GRMP_ERR_ARG,
                   "assemble_values: slot map or coefficient missing");
    int first[9];
    int k = 0;
    for (int rb = 0; rb < A->nblocks; ++rb) {
        first[rb] = k;
        for (auto &r : A->forms)
            if (r.f.row_block == rb) {
                fill_devform(A, r, hforms[k]);
                if (kappa && hforms[k].coef_cell) hforms[k].kappa = kappa;
                ++k;
            }
    }
    first[A->nblocks] = k;
    GRMP_HIP(ctx, hipMemcpyAsync(dforms, hforms, sizeof(DevForm) * k, hipMemcpyHostToDevice, st));
    const int B = asm_block_size(A->max_row);
    for (int rb = 0; rb < A->nblocks; ++rb) {
        FormPack P;
        P.f = dforms + first[rb];
        P.n = first[rb + 1] - first[rb];
        int64_t g0 = A->offs[rb], g1 = A->offs[rb + 1];
        if (g1 <= g0) continue;
        size_t shm = (size_t)B * (A->max_row > 0 ? A->max_row : 1) * sizeof(double);
        launch_assemble(A, A->spaces[rb], &hforms[first[rb]], P, g0, g1, B, shm, val64, val32, dinv64, dinv32, st);
        GRMP_HIP(ctx, hipGetLastError());
    }
    return GRMP_OK;
}

// ------------------------------------------------ batched assembly (many meshes)
// Many single-form matrices -- the levels of many served requests, or any
// set of meshes -- assembled by one launch: a persistent kernel whose
// workgroups walk a queue of (matrix, 256-row block) tiles and run, per tile,
// the serial row walk of k_assemble1s with that matrix's form.  Every row is
// summed in the order of the per-matrix kernels, so the values are the same
// bits.  Tiles go to workgroups in a fixed stride (no atomics).
struct AsmItemDev {
    DevForm F;
    const int32_t *rowptr;
    const uint8_t *diagpos, *fixed;
    double penalty;
    double *val64, *dinv64;
    float *dinv32;
    int64_t n;
};
struct AsmTile {
    int32_t item, blk;
};
#define ASMB_ROWS 256

__global__ __launch_bounds__(ASMB_ROWS) void k_assemble_queue(const AsmItemDev *__restrict__ items,
                                                               const AsmTile *__restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const AsmTile T = tiles[t];
        const AsmItemDev &I = items[T.item];
        const DevForm F = I.F;
        GRMP_DISPATCH(F, assemble_block([&](int64_t i, double *row) { assemble_row<KT>(F, i, row); }, (int64_t)T.blk,
                                        0, I.n, 0, I.rowptr, I.diagpos, I.fixed, I.penalty, I.val64, nullptr,
                                        I.dinv64, I.dinv32));
        __syncthreads();  // the next tile reuses the LDS rows
    }
}

int assemble_batch(grmp_ctx *ctx, hipStream_t st, const AsmBatchItem *items, int n)
{
    if (n <= 0) return GRMP_OK;
    std::vector<AsmItemDev> hi(n);
    std::vector<AsmTile> ht;
    int maxrow = 1;
    for (int k = 0; k < n; ++k) {
        const grmp_matrix *A = items[k].A;
        GRMP_CHECK(ctx, A && A->have_pattern && A->forms.size() == 1 && A->nblocks == 1 && items[k].val64 &&
                            items[k].dinv64 && items[k].dinv32,
                   GRMP_ERR_ARG, "batched assembly: item %d is not a single-form matrix with a pattern and outputs", k);
        AsmItemDev &d = hi[k];
        memset(&d, 0, sizeof(d));
        fill_devform(A, A->forms[0], d.F);
        if (items[k].kappa && d.F.coef_cell) d.F.kappa = items[k].kappa;
        GRMP_CHECK(ctx, d.F.kappa || !d.F.coef_cell, GRMP_ERR_ARG, "batched assembly: item %d has no coefficient", k);
        d.rowptr = A->rowptr;
        d.diagpos = A->diagpos;
        d.fixed = A->fixed;
        d.penalty = A->penalty;
        d.val64 = items[k].val64;
        d.dinv64 = items[k].dinv64;
        d.dinv32 = items[k].dinv32;
        d.n = A->n;
        maxrow = std::max(maxrow, A->max_row);
        for (int64_t b = 0; b * ASMB_ROWS < A->n; ++b) ht.push_back(AsmTile{k, (int32_t)b});
    }
    const size_t shm = (size_t)ASMB_ROWS * maxrow * sizeof(double);
    GRMP_CHECK(ctx, shm <= 64 * 1024, GRMP_ERR_UNSUPPORTED,
               "batched assembly: rows of %d entries exceed a tile's LDS (at most 32)", maxrow);
    const size_t bi = sizeof(AsmItemDev) * hi.size(), bt = sizeof(AsmTile) * ht.size();
    char *dbuf = nullptr;
    GRMP_HIP(ctx, hipMallocAsync((void **)&dbuf, bi + bt, st));
    hipError_t e = hipMemcpyAsync(dbuf, hi.data(), bi, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dbuf + bi, ht.data(), bt, hipMemcpyHostToDevice, st);
    const int ntiles = (int)ht.size();
    const int grid = std::min(ntiles, 2048);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_assemble_queue, dim3(grid), dim3(ASMB_ROWS), shm, st, (const AsmItemDev *)dbuf,
                           (const AsmTile *)(dbuf + bi), ntiles);
        e = hipGetLastError();
    }
    // the host arrays are the copies' sources: finish them before returning
    const hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFreeAsync(dbuf, st);
    GRMP_HIP(ctx, e);
    GRMP_HIP(ctx, e2);
    return GRMP_OK;
}

extern "C" int grmp_matrix_assemble_batch(int count, grmp_matrix *const *mats)
{
    GRMP_CHECK(nullptr, count >= 0 && (count == 0 || mats), GRMP_ERR_ARG, "grmp_matrix_assemble_batch: no matrices");
    if (count == 0) return GRMP_OK;
    grmp_ctx *ctx = mats[0]->ctx;
    std::vector<AsmBatchItem> items(count);
    for (int k = 0; k < count; ++k) {
        grmp_matrix *A = mats[k];
        GRMP_CHECK(ctx, A && A->ctx == ctx, GRMP_ERR_ARG, "grmp_matrix_assemble_batch: matrix %d of another context", k);
        if (!A->have_pattern) GRMP_TRY(grmp_matrix_build_pattern(A));
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    for (int k = 0; k < count; ++k) {
        grmp_matrix *A = mats[k];
        GRMP_CHECK(ctx, A->forms.size() == 1 && A->nblocks == 1, GRMP_ERR_UNSUPPORTED,
                   "grmp_matrix_assemble_batch: matrix %d has more than one form or block", k);
        if (!A->val64) GRMP_TRY(dalloc(ctx, &A->val64, A->nnz));
        if (!A->dinv64) GRMP_TRY(dalloc(ctx, &A->dinv64, A->n));
        if (!A->dinv32) GRMP_TRY(dalloc(ctx, &A->dinv32, A->n));
        items[k] = AsmBatchItem{A, nullptr, A->val64, A->dinv64, A->dinv32};
    }
    GRMP_TRY(assemble_batch(ctx, ctx->stream, items.data(), count));
    for (int k = 0; k < count; ++k) mats[k]->values |= GRMP_VALUES_FP64;
    return GRMP_OK;
}

extern "C" double grmp_matrix_last_assembly_seconds(const grmp_matrix *A) { return A ? A->last_asm_ms * 1e-3 : -1.0; }

extern "C" int grmp_matrix_set_tune(grmp_matrix *A, const grmp_asm_tune *t)
{
    if (!A) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    if (!t) {
        A->tune_set = false;
        return GRMP_OK;
    }
    GRMP_CHECK(ctx, t->variant >= 0 && t->variant <= 3 && t->rows_per_block >= 0 && t->rows_per_block <= 1024 &&
                        t->xcd_remap >= -1 && t->xcd_remap <= 1,
               GRMP_ERR_ARG, "grmp_matrix_set_tune: variant 0..3, rows_per_block 0..1024, xcd_remap -1..1");
    A->tune.variant = t->variant;
    A->tune.rows = t->rows_per_block;
    A->tune.remap = t->xcd_remap;
    A->tune_set = true;
    return GRMP_OK;
}

extern "C" int grmp_matrix_tune_key(const grmp_matrix *A, int *kind, int *nq, int *bucket)
{
    if (!A) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, A->have_pattern, GRMP_ERR_ARG, "grmp_matrix_tune_key: no pattern built");
    GRMP_CHECK(ctx, A->forms.size() == 1, GRMP_ERR_UNSUPPORTED,
               "grmp_matrix_tune_key: launch shapes are tuned for single-form matrices");
    if (kind) *kind = A->forms[0].kind;
    if (nq) *nq = A->forms[0].nq;
    if (bucket) *bucket = nnz_row_bucket(A->max_row);
    return GRMP_OK;
}

extern "C" int grmp_matrix_info(const grmp_matrix *A, int64_t *n, int64_t *nnz, int *max_row)
{
    if (!A) return GRMP_ERR_ARG;
    if (n) *n = A->n;
    if (nnz) *nnz = A->have_pattern ? A->nnz : -1;
    if (max_row) *max_row = A->max_row;
    return GRMP_OK;
}

// CSR of A -> CSC of A (1-based), i.e. CSR of A^T, built on the host.
extern "C" int grmp_matrix_export_csc(const grmp_matrix *A, int64_t *colptr, int64_t *rowval, double *nzval)
{
    if (!A || !colptr || !rowval) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, A->have_pattern, GRMP_ERR_ARG, "grmp_matrix_export_csc: no pattern built");
    std::lock_guard<std::mutex> g(ctx->mu);
    const int64_t n = A->n, nnz = A->nnz;
    std::vector<int32_t> rp(n + 1), cl(nnz);
    std::vector<double> vl(nzval ? nnz : 0);
    GRMP_HIP(ctx, hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipMemcpyAsync(cl.data(), A->cols, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, ctx->stream));
    if (nzval) {
        GRMP_CHECK(ctx, A->val64, GRMP_ERR_ARG, "grmp_matrix_export_csc: FP64 values not assembled");
        GRMP_HIP(ctx, hipMemcpyAsync(vl.data(), A->val64, sizeof(double) * nnz, hipMemcpyDeviceToHost, ctx->stream));
    }
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) cnt[cl[k] + 1]++;
    for (int64_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    for (int64_t j = 0; j <= n; ++j) colptr[j] = cnt[j] + 1;
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            int64_t p = cur[cl[k]]++;
            rowval[p] = i + 1;
            if (nzval) nzval[p] = vl[k];
        }
    return GRMP_OK;
}

// ------------------------------------------------------------ LinearForm
// b[off + i] += sum over cells of dof i (cell order) of localb * itemfactor,
// localb[a] = sum_q (sum_k f_k phi_k(a)) * w_q (src/assemblypatterns/linearform.jl:181-220).
struct LfDev {
    int fe, ncomp, nq, nloc, dim;
    double factor;
    double f[4];
    const double *coords;
    const int32_t *cells;
    int64_t nnodes, ncells;
    const int32_t *inc_ptr, *inc;
};

template <int FE, int NQ>
__device__ __forceinline__ double lf_local(const LfDev &L, int64_t e, int a)
{
    // cell volume only (the basis values are geometry independent)
    const int64_t nc = L.ncells, nn = L.nnodes;
    int32_t v0 = L.cells[e], v1 = L.cells[nc + e], v2 = L.cells[2 * nc + e];
    double A11 = L.coords[v1] - L.coords[v0], A12 = L.coords[v2] - L.coords[v0];
    double A21 = L.coords[nn + v1] - L.coords[nn + v0], A22 = L.coords[nn + v2] - L.coords[nn + v0];
    double det = A11 * A22 - A12 * A21;
    double vol = fabs(det) * 0.5;
    constexpr int NS = (FE == 1) ? 3 : 6;
    const int c = a / NS, s = a % NS;
    double lb = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x0, x1, w;
        qtri<NQ>(q, x0, x1, w);
        double phi = (FE == 1) ? p1tri_val(s, x0, x1) : p2tri_val(s, x0, x1);
        double t = 0.0;
        for (int k = 0; k < L.ncomp; ++k) t += L.f[k] * (k == c ? phi : 0.0);
        lb += t * w;
    }
    double itemfactor = L.factor * vol * 1.0;
    return lb * itemfactor;
}

// P1 on tetrahedra, order <= 1 (midpoint, weight 1, all four basis values 1/4)
__device__ __forceinline__ double lf_local_tet(const LfDev &L, int64_t e, int a)
{
    const int64_t nc = L.ncells, nn = L.nnodes;
    double X[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int32_t n = L.cells[v * nc + e];
        X[v][0] = L.coords[n];
        X[v][1] = L.coords[nn + n];
        X[v][2] = L.coords[2 * nn + n];
    }
    double A[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) A[r][k] = X[k + 1][r] - X[0][r];
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const double vol = fabs(det) / 6.0;
    const int c = a / 4;
    const double phi = 0.25;
    double t = 0.0;
    for (int k = 0; k < L.ncomp; ++k) t += L.f[k] * (k == c ? phi : 0.0);
    double lb = 0.0;
    lb += t * 1.0;
    const double itemfactor = L.factor * vol * 1.0;
    return lb * itemfactor;
}

__global__ void k_linearform_tet(LfDev L, double *__restrict__ b, int64_t ndofs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    double acc = b[i];
    for (int32_t t = L.inc_ptr[i]; t < L.inc_ptr[i + 1]; ++t) {
        const int32_t pk = L.inc[t];
        const int64_t e = pk / L.nloc;
        acc += lf_local_tet(L, e, pk - (int)e * L.nloc);
    }
    b[i] = acc;
}

template <int FE, int NQ>
__global__ void k_linearform(LfDev L, double *__restrict__ b, int64_t ndofs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    double acc = b[i];
    const int32_t t0 = L.inc_ptr[i], t1 = L.inc_ptr[i + 1];
    for (int32_t t = t0; t < t1; ++t) {
        const int32_t pk = L.inc[t];
        const int64_t e = pk / L.nloc;
        const int a = pk - (int)e * L.nloc;
        acc += lf_local<FE, NQ>(L, e, a);
    }
    b[i] = acc;
}

int linearform_launch(grmp_ctx *ctx, const grmp_space *s, const grmp_lform *form, double *bp, hipStream_t st)
{
    const grmp_mesh *m = s->mesh;
    int nq = quad_nq_tri(form->quadorder);
    const bool tet = m->dim == 3 && s->fe == GRMP_FE_H1P1 && form->quadorder <= 1;
    GRMP_CHECK(ctx, form->kind == GRMP_LF_CONST &&
                        (tet || (m->dim == 2 && (s->fe == GRMP_FE_H1P1 || s->fe == GRMP_FE_H1P2) && nq)),
               GRMP_ERR_UNSUPPORTED, "linear form kind %d on FE %d (dim %d, quadorder %d) not built", form->kind, s->fe,
               m->dim, form->quadorder);
    GRMP_CHECK(ctx, s->inc_ptr && s->inc && s->ncomp >= 1 && s->ncomp <= 4, GRMP_ERR_ARG,
               "linear form: space without dof-cell incidence");
    LfDev L;
    memset(&L, 0, sizeof(L));
    L.fe = s->fe;
    L.ncomp = s->ncomp;
    L.nq = nq;
    L.nloc = s->nloc;
    L.dim = m->dim;
    L.factor = form->factor;
    for (int k = 0; k < 4; ++k) L.f[k] = form->f[k];
    L.coords = m->coords;
    L.cells = m->cells;
    L.nnodes = m->nnodes;
    L.ncells = m->ncells;
    L.inc_ptr = s->inc_ptr;
    L.inc = s->inc;
    dim3 grid(nblk(s->ndofs, 256)), blk(256);
    if (tet) hipLaunchKernelGGL(k_linearform_tet, grid, blk, 0, st, L, bp, s->ndofs);
    else if (s->fe == GRMP_FE_H1P1 && nq == 1) hipLaunchKernelGGL((k_linearform<1, 1>), grid, blk, 0, st, L, bp, s->ndofs);
    else if (s->fe == GRMP_FE_H1P1) hipLaunchKernelGGL((k_linearform<1, 3>), grid, blk, 0, st, L, bp, s->ndofs);
    else if (nq == 1) hipLaunchKernelGGL((k_linearform<2, 1>), grid, blk, 0, st, L, bp, s->ndofs);
    else hipLaunchKernelGGL((k_linearform<2, 3>), grid, blk, 0, st, L, bp, s->ndofs);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

// ------------------------- polynomial data and the Newton right-hand side
// LinearForm(Identity, f) with f a polynomial of degree <= 4 per component
// (src/assemblypatterns/linearform.jl:181-220: per point the action's value
// times the test function, summed over the points, then times
// itemfactor = factor * vol), and the right-hand side full_assemble! adds for
// the Newton form (src/assemblypatterns/nonlinearform.jl:190-200, 218-224):
// (J [u, grad u] - kernel) . grad phi_a per point.  One thread per dof
// walking its cells in cell order.
#define LP_NCOEF 15
struct LpDev {
    int kind, fe, ncomp, per, nloc, nqt;
    double factor;
    double xq[RECO_MAXQ][2], wq[RECO_MAXQ];
    double g[2][LP_NCOEF];
    double nlc[3];
    const double *coords;
    const int32_t *cells, *celldofs;
    int64_t nnodes, ncells;
    const int32_t *inc_ptr, *inc;
    const double *uold;
};

__device__ __forceinline__ double lp_poly(const double *c, double x, double y)
{
    double s = 0.0;
    s += c[0];
    s += c[1] * x;
    s += c[2] * y;
    s += c[3] * (x * x);
    s += c[4] * (x * y);
    s += c[5] * (y * y);
    s += c[6] * (x * x * x);
    s += c[7] * (x * x * y);
    s += c[8] * (x * y * y);
    s += c[9] * (y * y * y);
    s += c[10] * (x * x * x * x);
    s += c[11] * (x * x * x * y);
    s += c[12] * (x * x * y * y);
    s += c[13] * (x * y * y * y);
    s += c[14] * (y * y * y * y);
    return s;
}

template <int FE>
__device__ __forceinline__ double lp_local(const LpDev &L, int64_t e, int a)
{
    constexpr int ND = (FE == 1) ? 3 : 6;
    const int64_t nc = L.ncells, nn = L.nnodes;
    const int32_t v0 = L.cells[e], v1 = L.cells[nc + e], v2 = L.cells[2 * nc + e];
    const double X0 = L.coords[v0], Y0 = L.coords[nn + v0];
    const double A11 = L.coords[v1] - X0, A12 = L.coords[v2] - X0;
    const double A21 = L.coords[nn + v1] - Y0, A22 = L.coords[nn + v2] - Y0;
    const double det = A11 * A22 - A12 * A21;
    const double vol = fabs(det) * 0.5;
    double lb = 0.0;
    if (L.kind == GRMP_LF_POLY) {
        const int c = a / L.per, s = a - c * L.per;
        for (int q = 0; q < L.nqt; ++q) {
            const double x0 = L.xq[q][0], x1 = L.xq[q][1];
            double xp = 0.0, yp = 0.0;
            xp += A11 * x0;
            xp += A12 * x1;
            xp += X0;
            yp += A21 * x0;
            yp += A22 * x1;
            yp += Y0;
            const double phi = FE == 1 ? p1tri_val(s, x0, x1) : p2tri_val(s, x0, x1);
            double temp = 0.0;
            for (int k = 0; k < L.ncomp; ++k) temp += lp_poly(L.g[k], xp, yp) * (k == c ? phi : 0.0);
            lb += temp * L.wq[q];
        }
    } else {
        // Newton right-hand side: A^{-T} as the matrix kernels form it
        double M[2][2];
        M[0][0] = A22 / det;
        M[0][1] = -A21 / det;
        M[1][0] = -A12 / det;
        M[1][1] = A11 / det;
        double uc[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) uc[d] = L.uold[L.celldofs[(int64_t)d * nc + e]];
        for (int q = 0; q < L.nqt; ++q) {
            const double x0 = L.xq[q][0], x1 = L.xq[q][1];
            double in[3], J[2][3], val[2];
            nl_point(uc, ND, FE, M, x0, x1, in);
            nl_jac(L.nlc, in, J, val);
            double ga[2];
            grad_tri<FE>(M, a, x0, x1, ga);
            double temp = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double ar = 0.0;
                ar += J[k][0] * in[0];
                ar += J[k][1] * in[1];
                ar += J[k][2] * in[2];
                temp += (ar - val[k]) * ga[k];
            }
            lb += temp * L.wq[q];
        }
    }
    return lb * (L.factor * vol * 1.0);
}

template <int FE>
__global__ void k_linearform_lp(LpDev L, double *__restrict__ b, int64_t ndofs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    double acc = b[i];
    for (int32_t t = L.inc_ptr[i]; t < L.inc_ptr[i + 1]; ++t) {
        const int32_t pk = L.inc[t];
        const int64_t e = pk / L.nloc;
        acc += lp_local<FE>(L, e, pk - (int)e * L.nloc);
    }
    b[i] = acc;
}

static int linearform_lp_launch(grmp_ctx *ctx, const grmp_space *s, const grmp_lform *form, double *bp, hipStream_t st)
{
    const grmp_mesh *m = s->mesh;
    GRMP_CHECK(ctx, m->dim == 2 && m->nloc == 3 && (s->fe == GRMP_FE_H1P1 || s->fe == GRMP_FE_H1P2), GRMP_ERR_UNSUPPORTED,
               "linear form kind %d: P1/P2 on triangles only", form->kind);
    GRMP_CHECK(ctx, s->inc_ptr && s->inc, GRMP_ERR_ARG, "linear form: space without dof-cell incidence");
    LpDev L;
    memset(&L, 0, sizeof(L));
    L.kind = form->kind;
    L.fe = s->fe;
    L.ncomp = s->ncomp;
    L.per = s->fe == GRMP_FE_H1P2 ? 6 : 3;
    L.nloc = s->nloc;
    L.nqt = tri_rule(form->quadorder, L.xq, L.wq);
    GRMP_CHECK(ctx, L.nqt > 0, GRMP_ERR_UNSUPPORTED, "linear form: quadrature order %d too high", form->quadorder);
    L.factor = form->factor;
    if (form->kind == GRMP_LF_POLY) {
        GRMP_CHECK(ctx, s->ncomp <= 2 && form->coef && form->ncoef >= 1 && form->ncoef <= LP_NCOEF, GRMP_ERR_ARG,
                   "GRMP_LF_POLY: 1..%d coefficients per component, at most 2 components", LP_NCOEF);
        for (int k = 0; k < s->ncomp; ++k)
            for (int j = 0; j < form->ncoef; ++j) L.g[k][j] = form->coef[k * form->ncoef + j];
    } else {
        GRMP_CHECK(ctx, s->ncomp == 1 && form->current && form->current_offset >= 0 &&
                            form->current_offset + s->ndofs <= form->current->n,
                   GRMP_ERR_ARG, "GRMP_LF_NEWTON: a scalar space and its iterate vector are needed");
        L.uold = form->current->d + form->current_offset;
        for (int k = 0; k < 3; ++k) L.nlc[k] = form->f[k];
    }
    L.coords = m->coords;
    L.cells = m->cells;
    L.celldofs = s->celldofs;
    L.nnodes = m->nnodes;
    L.ncells = m->ncells;
    L.inc_ptr = s->inc_ptr;
    L.inc = s->inc;
    const dim3 grid(nblk(s->ndofs, 256)), blk(256);
    if (s->fe == GRMP_FE_H1P1) hipLaunchKernelGGL(k_linearform_lp<1>, grid, blk, 0, st, L, bp, s->ndofs);
    else hipLaunchKernelGGL(k_linearform_lp<2>, grid, blk, 0, st, L, bp, s->ndofs);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

// ------------------------------------------ reconstructed Stokes right-hand side
// LinearForm(ReconstructionIdentity{HDIVRT0{2}}, f) on H1BR{2}
// (src/feevaluator_h1.jl:342-381, src/reconstructions.jl:381-403): the test
// function is replaced by its RT0 reconstruction sum_f rc[a][f] Piola(RT0_f),
// rc = |F|/2 n_k(F) for the node dofs of face F, |F| for its bubble.

#define RECO_MAXMODES 16
struct RecoDev {
    int nq, K;
    double xq[RECO_MAXQ][2], wq[RECO_MAXQ];
    double modes[2 * RECO_MAXMODES][4];
    double poly[4];
    double factor;
    const double *coords;
    const int32_t *cells;
    const int8_t *fsign;
    int64_t nnodes, ncells;
    const int32_t *inc_ptr, *inc;
};

// symmetric eigenpairs by cyclic Jacobi rotations (n <= 8), ascending
static void jacobi_eigen(int n, double *A, double *ev, double *V)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
        if (off < 1e-60) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double x = A[k * n + p], y = A[k * n + q];
                    A[k * n + p] = c * x - sn * y;
                    A[k * n + q] = sn * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = A[p * n + k], y = A[q * n + k];
                    A[p * n + k] = c * x - sn * y;
                    A[q * n + k] = sn * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = V[k * n + p], y = V[k * n + q];
                    V[k * n + p] = c * x - sn * y;
                    V[k * n + q] = sn * x + c * y;
                }
            }
    }
    for (int i = 0; i < n; ++i) ev[i] = A[i * n + i];
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (ev[j] < ev[i]) {
                std::swap(ev[i], ev[j]);
                for (int k = 0; k < n; ++k) std::swap(V[k * n + i], V[k * n + j]);
            }
}

// Triangle rule of the given order: orders <= 2 as qtri, otherwise the
// Stroud conical product of Gauss-Legendre and Gauss-Jacobi(1,0) points
// (src/quadrature.jl:528-562), both from their Jacobi matrices (Golub-Welsch).
static int tri_rule(int order, double (*xq)[2], double *wq)
{
    if (order <= 1) {
        xq[0][0] = xq[0][1] = 1.0 / 3.0;
        wq[0] = 1.0;
        return 1;
    }
    if (order == 2) {
        const double pts[3][2] = {{0.5, 0.5}, {0.0, 0.5}, {0.5, 0.0}};
        for (int q = 0; q < 3; ++q) {
            xq[q][0] = pts[q][0];
            xq[q][1] = pts[q][1];
            wq[q] = 1.0 / 3.0;
        }
        return 3;
    }
    const int n = order / 2 + 1;
    if (n * n > RECO_MAXQ) return 0;
    double J[64], ev[8], V[64], r[8], a[8], sj[8], bj[8];
    for (int i = 0; i < n * n; ++i) J[i] = 0.0;
    for (int k = 1; k < n; ++k) J[(k - 1) * n + k] = J[k * n + k - 1] = (double)k / std::sqrt(4.0 * k * k - 1.0);
    jacobi_eigen(n, J, ev, V);
    for (int i = 0; i < n; ++i) {
        r[i] = 0.5 * ev[i] + 0.5;
        a[i] = 0.5 * (2.0 * V[i] * V[i]);
    }
    for (int i = 0; i < n * n; ++i) J[i] = 0.0;
    for (int k = 1; k <= n; ++k) J[(k - 1) * n + k - 1] = -1.0 / (4.0 * k * k - 1.0);
    for (int k = 1; k < n; ++k)
        J[(k - 1) * n + k] = J[k * n + k - 1] = std::sqrt((double)(k + 1) * k) / (2.0 * (k + 1) - 1.0);
    jacobi_eigen(n, J, ev, V);
    for (int i = 0; i < n; ++i) {
        sj[i] = 0.5 * ev[i] + 0.5;
        bj[i] = 0.5 * (2.0 * V[i] * V[i]);
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            xq[j * n + i][0] = sj[j];
            xq[j * n + i][1] = r[i] * (1.0 - sj[j]);
            wq[j * n + i] = a[i] * bj[j];
        }
    return n * n;
}

__device__ __forceinline__ void stokes_force(const RecoDev &R, double x, double y, double &fx, double &fy)
{
    const double tp = 6.283185307179586;
    fx = 0.0;
    fy = 0.0;
    for (int m = 0; m < R.K; ++m) {
        const double *md = R.modes[m];
        const double cs = md[0] * tp * cos(tp * (md[1] * x + md[2] * y) + md[3]);
        fx += cs * md[1];
        fy += cs * md[2];
    }
    for (int m = 0; m < R.K; ++m) {
        const double *md = R.modes[R.K + m];
        const double cs = md[0] * tp * cos(tp * (md[1] * x + md[2] * y) + md[3]);
        fx += cs * md[2];
        fy -= cs * md[1];
    }
    fx += 3.0 * R.poly[0] * x * x + R.poly[1] * y * y;
    fy += 2.0 * R.poly[1] * x * y + 2.0 * R.poly[2] * y;
}

__device__ double reco_local(const RecoDev &R, int64_t e, int a)
{
    const int64_t nc = R.ncells, nn = R.nnodes;
    int32_t v[3];
    double X[3], Y[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = R.cells[k * nc + e];
        X[k] = R.coords[v[k]];
        Y[k] = R.coords[nn + v[k]];
    }
    const double A11 = X[1] - X[0], A12 = X[2] - X[0], A21 = Y[1] - Y[0], A22 = Y[2] - Y[0];
    const double det = A11 * A22 - A12 * A21;
    const double vol = fabs(det) * 0.5;
    double sgn[3], rc[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const bool own = R.fsign[f * nc + e] > 0;
        sgn[f] = own ? 1.0 : -1.0;
        const int ia = own ? f : (f + 1) % 3, ib = own ? (f + 1) % 3 : f;
        const double tx = X[ib] - X[ia], ty = Y[ib] - Y[ia];
        const double len = sqrt(tx * tx + ty * ty);
        const double nk[2] = {ty / len, -tx / len};
        if (a < 6) {
            const int node = a % 3, k = a / 3;
            const bool on = node == f || node == (f + 1) % 3;
            rc[f] = on ? 0.5 * len * nk[k] : 0.0;
        } else {
            rc[f] = (a - 6 == f) ? len : 0.0;
        }
    }
    double lb = 0.0;
    for (int q = 0; q < R.nq; ++q) {
        const double x0 = R.xq[q][0], x1 = R.xq[q][1];
        const double phi[3][2] = {{x0, x1 - 1.0}, {x0, x1}, {x0 - 1.0, x1}};
        double cv0 = 0.0, cv1 = 0.0;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if (rc[f] == 0.0) continue;
            double t0 = 0.0, t1 = 0.0;
            t0 += A11 * phi[f][0];
            t0 += A12 * phi[f][1];
            t1 += A21 * phi[f][0];
            t1 += A22 * phi[f][1];
            const double sd = sgn[f] / det;
            t0 *= sd;
            t1 *= sd;
            cv0 += rc[f] * t0;
            cv1 += rc[f] * t1;
        }
        const double xp = X[0] + A11 * x0 + A12 * x1, yp = Y[0] + A21 * x0 + A22 * x1;
        double fx, fy;
        stokes_force(R, xp, yp, fx, fy);
        double t = 0.0;
        t += fx * cv0;
        t += fy * cv1;
        lb += t * R.wq[q];
    }
    return lb * (R.factor * vol * 1.0);
}

__global__ void k_linearform_reco(RecoDev R, double *__restrict__ b, int64_t ndofs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    double acc = b[i];
    for (int32_t t = R.inc_ptr[i]; t < R.inc_ptr[i + 1]; ++t) {
        const int32_t pk = R.inc[t];
        const int64_t e = pk / 9;
        acc += reco_local(R, e, pk - (int)e * 9);
    }
    b[i] = acc;
}

static int linearform_reco_launch(grmp_ctx *ctx, const grmp_space *s, const grmp_lform *form, double *bp, hipStream_t st)
{
    const grmp_mesh *m = s->mesh;
    GRMP_CHECK(ctx, s->fe == GRMP_FE_H1BR && s->nloc == 9 && m->dim == 2 && m->fsign, GRMP_ERR_UNSUPPORTED,
               "ReconstructionIdentity{HDIVRT0{2}} needs H1BR{2} on a mesh with CellFaceSigns");
    GRMP_CHECK(ctx, s->inc_ptr && s->inc, GRMP_ERR_ARG, "RECO_RT0 form: the space has no dof-cell incidence");
    GRMP_CHECK(ctx, form->nmodes >= 0 && form->nmodes <= RECO_MAXMODES && (form->nmodes == 0 || form->modes),
               GRMP_ERR_ARG, "RECO_RT0 form: at most %d modes", RECO_MAXMODES);
    RecoDev R;
    memset(&R, 0, sizeof(R));
    R.nq = tri_rule(form->quadorder, R.xq, R.wq);
    GRMP_CHECK(ctx, R.nq > 0, GRMP_ERR_UNSUPPORTED, "RECO_RT0 form: quadrature order %d too high", form->quadorder);
    R.K = form->nmodes;
    for (int m2 = 0; m2 < 2 * R.K; ++m2)
        for (int k = 0; k < 4; ++k) R.modes[m2][k] = form->modes[4 * m2 + k];
    for (int k = 0; k < 4; ++k) R.poly[k] = form->poly[k];
    R.factor = form->factor;
    R.coords = m->coords;
    R.cells = m->cells;
    R.fsign = m->fsign;
    R.nnodes = m->nnodes;
    R.ncells = m->ncells;
    R.inc_ptr = s->inc_ptr;
    R.inc = s->inc;
    hipLaunchKernelGGL(k_linearform_reco, dim3(nblk(s->ndofs, 128)), dim3(128), 0, st, R, bp, s->ndofs);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

extern "C" int grmp_assemble_linearform(grmp_ctx *ctx, const grmp_space *s, const grmp_lform *form, grmp_vector *b,
                                        int64_t offset)
{
    GRMP_CHECK(ctx, ctx && s && form && b, GRMP_ERR_ARG, "grmp_assemble_linearform: null argument");
    GRMP_CHECK(ctx, offset >= 0 && offset + s->ndofs <= b->n, GRMP_ERR_ARG, "grmp_assemble_linearform: vector too short");
    std::lock_guard<std::mutex> g(ctx->mu);
    if (form->kind == GRMP_LF_RECO_RT0) return linearform_reco_launch(ctx, s, form, b->d + offset, ctx->stream);
    if (form->kind == GRMP_LF_POLY || form->kind == GRMP_LF_NEWTON)
        return linearform_lp_launch(ctx, s, form, b->d + offset, ctx->stream);
    return linearform_launch(ctx, s, form, b->d + offset, ctx->stream);
}

// --------------------------------------------------------- ItemIntegrator
// per cell: out[k, c] = sum_q (sum_dofs u_dof * phi_k(dof)) * w_q * vol
// (NoAction branch of src/assemblypatterns/itemintegrator.jl:276-283)
template <int FE, int NQ>
__global__ void k_integrate_identity(LfDev L, const int32_t *__restrict__ dofs, const double *__restrict__ u,
                                     double *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.ncells) return;
    const int64_t nc = L.ncells, nn = L.nnodes;
    int32_t v0 = L.cells[e], v1 = L.cells[nc + e], v2 = L.cells[2 * nc + e];
    double A11 = L.coords[v1] - L.coords[v0], A12 = L.coords[v2] - L.coords[v0];
    double A21 = L.coords[nn + v1] - L.coords[nn + v0], A22 = L.coords[nn + v2] - L.coords[nn + v0];
    double vol = fabs(A11 * A22 - A12 * A21) * 0.5;
    constexpr int NS = (FE == 1) ? 3 : 6;
    for (int k = 0; k < L.ncomp; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x0, x1, w;
            qtri<NQ>(q, x0, x1, w);
            double in = 0.0;
            for (int c = 0; c < L.ncomp; ++c)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    double phi = (FE == 1) ? p1tri_val(s, x0, x1) : p2tri_val(s, x0, x1);
                    double uv = u[dofs[(int64_t)(c * NS + s) * nc + e]];
                    in += uv * (c == k ? phi : 0.0);
                }
            acc += in * w * vol;
        }
        out[e * L.ncomp + k] = acc;
    }
}

// P0 (NoAction, one quadrature point of weight 1): out[c] = (0 + u_c*1) * 1 * vol
__global__ void k_integrate_p0(LfDev L, const double *__restrict__ u, double *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.ncells) return;
    const int64_t nc = L.ncells, nn = L.nnodes;
    int32_t v0 = L.cells[e], v1 = L.cells[nc + e], v2 = L.cells[2 * nc + e];
    double A11 = L.coords[v1] - L.coords[v0], A12 = L.coords[v2] - L.coords[v0];
    double A21 = L.coords[nn + v1] - L.coords[nn + v0], A22 = L.coords[nn + v2] - L.coords[nn + v0];
    double vol = fabs(A11 * A22 - A12 * A21) * 0.5;
    double in = 0.0;
    in += u[e] * 1.0;
    double acc = 0.0;
    acc += in * 1.0 * vol;
    out[e] = acc;
}

extern "C" int grmp_integrate_identity(grmp_ctx *ctx, const grmp_space *s, int quadorder, const grmp_vector *u,
                                       int64_t offset, double *out)
{
    GRMP_CHECK(ctx, ctx && s && u && out, GRMP_ERR_ARG, "grmp_integrate_identity: null argument");
    GRMP_CHECK(ctx, offset >= 0 && offset + s->ndofs <= u->n, GRMP_ERR_ARG, "grmp_integrate_identity: vector too short");
    const grmp_mesh *m = s->mesh;
    int nq = quad_nq_tri(quadorder);
    GRMP_CHECK(ctx, m->dim == 2 && (s->fe == GRMP_FE_H1P1 || s->fe == GRMP_FE_H1P2 || s->fe == GRMP_FE_L2P0) && nq,
               GRMP_ERR_UNSUPPORTED, "grmp_integrate_identity: FE %d in %dD with quadorder %d not built", s->fe, m->dim,
               quadorder);
    std::lock_guard<std::mutex> g(ctx->mu);
    LfDev L;
    memset(&L, 0, sizeof(L));
    L.ncomp = s->ncomp;
    L.coords = m->coords;
    L.cells = m->cells;
    L.nnodes = m->nnodes;
    L.ncells = m->ncells;
    double *d = nullptr;
    GRMP_TRY(dalloc(ctx, &d, m->ncells * s->ncomp));
    dim3 grid(nblk(m->ncells, 256)), blk(256);
    const double *up = u->d + offset;
    if (s->fe == GRMP_FE_L2P0) hipLaunchKernelGGL(k_integrate_p0, grid, blk, 0, ctx->stream, L, up, d);
    else if (s->fe == GRMP_FE_H1P1 && nq == 1) hipLaunchKernelGGL((k_integrate_identity<1, 1>), grid, blk, 0, ctx->stream, L, s->celldofs, up, d);
    else if (s->fe == GRMP_FE_H1P1) hipLaunchKernelGGL((k_integrate_identity<1, 3>), grid, blk, 0, ctx->stream, L, s->celldofs, up, d);
    else if (nq == 1) hipLaunchKernelGGL((k_integrate_identity<2, 1>), grid, blk, 0, ctx->stream, L, s->celldofs, up, d);
    else hipLaunchKernelGGL((k_integrate_identity<2, 3>), grid, blk, 0, ctx->stream, L, s->celldofs, up, d);
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_HIP(ctx, hipMemcpyAsync(out, d, sizeof(double) * m->ncells * s->ncomp, hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, d);
    return GRMP_OK;
}

// ----------------------------------------------------- L2ErrorIntegrator
// L2ErrorIntegrator(g, Identity; quadorder, factor) and L2NormIntegrator
// (src/assemblypatterns/itemintegrator.jl:40-95): evaluate! (:243-307) with
// the action L2error_function, per cell
//   out[c] = sum_q (sum_k (g_k(x_q) - u_h,k(x_q) * factor)^2) * w_q * vol,
// u_h,k(x_q) = sum_dofs u_dof * cvals[k, dof, q] (eval_febe!, src/feevaluator.jl:445-452),
// x_q = x_1 + A xref_q.  g is a polynomial of degree <= 3 per component
// (ncoef 0: g = 0, the L2 norm).  P1, P2 (1 or 2 components), H1BR{2}
// (Identity with coefficients: face bubbles times the face normal,
// src/feevaluator_h1.jl:31-42) and L2P0 on triangles.
#define L2E_MAXQ RECO_MAXQ
struct L2eDev {
    double xq[L2E_MAXQ][2], wq[L2E_MAXQ];
    double g[2][10];
    double factor;
    int nq, fe, ncomp, per;
    const double *coords;
    const int32_t *cells, *celldofs;
    const int8_t *fsign;
    int64_t nnodes, ncells;
};

__device__ __forceinline__ double l2e_poly(const double *c, double x, double y)
{
    double s = 0.0;
    s += c[0];
    s += c[1] * x;
    s += c[2] * y;
    s += c[3] * (x * x);
    s += c[4] * (x * y);
    s += c[5] * (y * y);
    s += c[6] * (x * x * x);
    s += c[7] * (x * x * y);
    s += c[8] * (x * y * y);
    s += c[9] * (y * y * y);
    return s;
}

// value of scalar shape function sf of the element at (x0, x1)
__device__ __forceinline__ double l2e_shape(int fe, int sf, double x0, double x1)
{
    if (fe == GRMP_FE_H1P1) return p1tri_val(sf, x0, x1);
    if (fe == GRMP_FE_H1P2) return p2tri_val(sf, x0, x1);
    // H1BR{2}: P1 nodes, then bubbles 6 x l, 6 y x, 6 l y (src/fedefs/h1v_br.jl:117-130)
    double l = 1.0;
    l -= x0;
    l -= x1;
    if (sf < 3) return sf == 0 ? l : (sf == 1 ? x0 : x1);
    if (sf == 3) return 6.0 * x0 * l;
    if (sf == 4) return 6.0 * x1 * x0;
    return 6.0 * l * x1;
}

__global__ void k_l2error(L2eDev L, const double *__restrict__ u, double *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.ncells) return;
    const int64_t nc = L.ncells, nn = L.nnodes;
    const int32_t v0 = L.cells[e], v1 = L.cells[nc + e], v2 = L.cells[2 * nc + e];
    const double X0 = L.coords[v0], Y0 = L.coords[nn + v0];
    const double A11 = L.coords[v1] - X0, A12 = L.coords[v2] - X0;
    const double A21 = L.coords[nn + v1] - Y0, A22 = L.coords[nn + v2] - Y0;
    const double vol = fabs(A11 * A22 - A12 * A21) * 0.5;
    double n[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    if (L.fe == GRMP_FE_H1BR) {
        const int32_t vv[3] = {v0, v1, v2};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const bool own = L.fsign[k * nc + e] > 0;
            const int32_t a = own ? vv[k] : vv[(k + 1) % 3], b = own ? vv[(k + 1) % 3] : vv[k];
            const double tx = L.coords[b] - L.coords[a], ty = L.coords[nn + b] - L.coords[nn + a];
            const double len = sqrt(tx * tx + ty * ty);
            n[k][0] = ty / len;
            n[k][1] = -tx / len;
        }
    }
    double acc = 0.0;
    for (int q = 0; q < L.nq; ++q) {
        const double x0 = L.xq[q][0], x1 = L.xq[q][1];
        double in[2] = {0.0, 0.0};
        if (L.fe == GRMP_FE_L2P0) {
            in[0] += u[e] * 1.0;
        } else if (L.fe == GRMP_FE_H1BR) {
            for (int dof = 0; dof < 9; ++dof) {
                const double ud = u[L.celldofs[(int64_t)dof * nc + e]];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double cv;
                    if (dof < 6) cv = (dof / 3 == k) ? l2e_shape(L.fe, dof % 3, x0, x1) * 1.0 : 0.0;
                    else cv = l2e_shape(L.fe, dof - 3, x0, x1) * n[dof - 6][k];
                    in[k] += ud * cv * 1.0;
                }
            }
        } else {
            for (int c = 0; c < L.ncomp; ++c)
                for (int s = 0; s < L.per; ++s) {
                    const double ud = u[L.celldofs[(int64_t)(c * L.per + s) * nc + e]];
                    for (int k = 0; k < L.ncomp; ++k) in[k] += ud * (c == k ? l2e_shape(L.fe, s, x0, x1) : 0.0) * 1.0;
                }
        }
        double xp = 0.0, yp = 0.0;
        xp += A11 * x0;
        xp += A12 * x1;
        xp += X0;
        yp += A21 * x0;
        yp += A22 * x1;
        yp += Y0;
        double r = 0.0;
        for (int k = 0; k < L.ncomp; ++k) {
            double val = l2e_poly(L.g[k], xp, yp);
            val -= in[k] * L.factor;
            r += val * val;
        }
        acc += r * L.wq[q] * vol;
    }
    out[e] = acc;
}

// fixed-order sum: each block sums its 1024 entries by a fixed tree, then one
// block sums the block partials the same way (deterministic for a given n)
__device__ __forceinline__ double l2e_block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    return s;
}

__global__ __launch_bounds__(1024) void k_sum_blocks(const double *__restrict__ x, int64_t n, double *__restrict__ part)
{
    __shared__ double red[16];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    const double s = l2e_block_sum(i < n ? x[i] : 0.0, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_sum_final(const double *__restrict__ part, int64_t np, double *__restrict__ out)
{
    __shared__ double red[16];
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += 1024) v += part[i];
    const double s = l2e_block_sum(v, red);
    if (threadIdx.x == 0) out[0] = s;
}

static int l2e_setup(grmp_ctx *ctx, const grmp_space *s, const grmp_vector *u, int64_t offset, const double *coef,
                     int ncoef, int quadorder, double factor, L2eDev &L)
{
    GRMP_CHECK(ctx, ctx && s && u, GRMP_ERR_ARG, "grmp_integrate_l2error: null argument");
    GRMP_CHECK(ctx, offset >= 0 && offset + s->ndofs <= u->n, GRMP_ERR_ARG, "grmp_integrate_l2error: vector too short");
    GRMP_CHECK(ctx, ncoef >= 0 && ncoef <= 10 && (ncoef == 0 || coef), GRMP_ERR_ARG,
               "grmp_integrate_l2error: 0..10 polynomial coefficients per component");
    const grmp_mesh *m = s->mesh;
    GRMP_CHECK(ctx, m->dim == 2 && m->nloc == 3, GRMP_ERR_UNSUPPORTED, "grmp_integrate_l2error: triangle meshes only");
    GRMP_CHECK(ctx, ((s->fe == GRMP_FE_H1P1 || s->fe == GRMP_FE_H1P2) && s->ncomp <= 2) || s->fe == GRMP_FE_L2P0 ||
                        (s->fe == GRMP_FE_H1BR && m->fsign),
               GRMP_ERR_UNSUPPORTED, "grmp_integrate_l2error: H1P1/H1P2 (1-2 components), H1BR{2}, L2P0 only");
    memset(&L, 0, sizeof(L));
    L.nq = tri_rule(quadorder, L.xq, L.wq);
    GRMP_CHECK(ctx, L.nq > 0, GRMP_ERR_UNSUPPORTED, "grmp_integrate_l2error: quadrature order %d too high", quadorder);
    L.fe = s->fe;
    L.ncomp = s->ncomp;
    L.per = s->fe == GRMP_FE_H1P2 ? 6 : 3;
    for (int k = 0; k < L.ncomp; ++k)
        for (int j = 0; j < ncoef; ++j) L.g[k][j] = coef[k * ncoef + j];
    L.factor = factor;
    L.coords = m->coords;
    L.cells = m->cells;
    L.celldofs = s->celldofs;
    L.fsign = m->fsign;
    L.nnodes = m->nnodes;
    L.ncells = m->ncells;
    return GRMP_OK;
}

// per-cell values into d[0..nc), block partials into d[nc..nc+nb), the total
// into d[nc+nb], then both copied to the host; every element of d is written
// before it is read, so d needs no clearing
static int l2e_run(grmp_ctx *ctx, const L2eDev &L, const double *u, double *d, double *cellvals, double *total)
{
    const int64_t nc = L.ncells, nb = (nc + 1023) / 1024;
    hipLaunchKernelGGL(k_l2error, dim3(nblk(nc, 256)), dim3(256), 0, ctx->stream, L, u, d);
    hipLaunchKernelGGL(k_sum_blocks, dim3((unsigned)nb), dim3(1024), 0, ctx->stream, d, nc, d + nc);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(1024), 0, ctx->stream, d + nc, nb, d + nc + nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && cellvals)
        e = hipMemcpyAsync(cellvals, d, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && total)
        e = hipMemcpyAsync(total, d + nc + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    GRMP_HIP(ctx, e);
    return GRMP_OK;
}

extern "C" int grmp_integrate_l2error(grmp_ctx *ctx, const grmp_space *s, const grmp_vector *u, int64_t offset,
                                      const double *coef, int ncoef, int quadorder, double factor, double *cellvals,
                                      double *total)
{
    GRMP_CHECK(ctx, ctx && (cellvals || total), GRMP_ERR_ARG, "grmp_integrate_l2error: no output");
    L2eDev L;
    GRMP_TRY(l2e_setup(ctx, s, u, offset, coef, ncoef, quadorder, factor, L));
    std::lock_guard<std::mutex> g(ctx->mu);
    const int64_t nc = L.ncells, nb = (nc + 1023) / 1024;
    double *d = nullptr;
    GRMP_TRY(dalloc(ctx, &d, nc + nb + 1));
    const int rc = l2e_run(ctx, L, u->d + offset, d, cellvals, total);
    dfree(ctx, d);
    return rc;
}

// ------------------------------- L2ErrorIntegrator without device allocation
// grmp_integrate_l2error_noalloc keeps its per-cell values and partials in
// this file-scope array instead of allocating them.  Each GPU holds its own
// copy, at its own address, until the process ends.  The host never uses the
// array's name as an address: the per-device address comes from
// hipGetSymbolAddress, taken once per device on the entry point's first call
// there with that device current.
#define L2E_NA_MAXCELLS (1ll << 22)
#define L2E_NA_WORDS (L2E_NA_MAXCELLS + (L2E_NA_MAXCELLS + 1023) / 1024 + 1)
static __device__ __align__(256) double scratch[L2E_NA_WORDS];

#define L2E_NA_MAXDEV 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[L2E_NA_MAXDEV];

static int l2e_scratch(grmp_ctx *ctx, double **out)
{
    const int dev = ctx->device;
    GRMP_CHECK(ctx, dev >= 0 && dev < L2E_NA_MAXDEV, GRMP_ERR_UNSUPPORTED, "no-allocation scratch: device %d", dev);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int prev = 0;
        GRMP_HIP(ctx, hipGetDevice(&prev));
        GRMP_HIP(ctx, hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        GRMP_HIP(ctx, e);
        GRMP_CHECK(ctx, p != nullptr, GRMP_ERR_HIP, "no-allocation scratch: no device address");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return GRMP_OK;
}

extern "C" int grmp_integrate_l2error_noalloc(grmp_ctx *ctx, const grmp_space *s, const grmp_vector *u,
                                              int64_t offset, const double *coef, int ncoef, int quadorder,
                                              double factor, double *cellvals, double *total)
{
    GRMP_CHECK(ctx, ctx && (cellvals || total), GRMP_ERR_ARG, "grmp_integrate_l2error_noalloc: no output");
    L2eDev L;
    GRMP_TRY(l2e_setup(ctx, s, u, offset, coef, ncoef, quadorder, factor, L));
    GRMP_CHECK(ctx, L.ncells <= L2E_NA_MAXCELLS, GRMP_ERR_ARG,
               "grmp_integrate_l2error_noalloc: %lld cells exceed the scratch limit of %lld cells",
               (long long)L.ncells, (long long)L2E_NA_MAXCELLS);
    double *d = nullptr;
    GRMP_TRY(l2e_scratch(ctx, &d));
    std::lock_guard<std::mutex> g(ctx->mu);
    return l2e_run(ctx, L, u->d + offset, d, cellvals, total);
}

// ------------------------------------------------------------ CellVolumes
__global__ void k_cellvolumes(const double *coords, const int32_t *cells, int64_t nn, int64_t nc, int dim, double *out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc) return;
    DevForm F;
    memset(&F, 0, sizeof(F));
    F.coords = coords;
    F.cells = cells;
    F.nnodes = nn;
    F.ncells = nc;
    double vol;
    if (dim == 2) {
        double M[2][2];
        trafo_tri(F, e, M, vol);
    } else {
        double M[3][3];
        trafo_tet(F, e, M, vol);
    }
    out[e] = vol;
}

extern "C" int grmp_mesh_cellvolumes(const grmp_mesh *m, double *out)
{
    if (!m || !out) return GRMP_ERR_ARG;
    grmp_ctx *ctx = m->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    double *d = nullptr;
    GRMP_TRY(dalloc(ctx, &d, m->ncells));
    hipLaunchKernelGGL(k_cellvolumes, dim3(nblk(m->ncells, 256)), dim3(256), 0, ctx->stream, m->coords, m->cells,
                       m->nnodes, m->ncells, m->dim, d);
    GRMP_HIP(ctx, hipGetLastError());
    GRMP_HIP(ctx, hipMemcpyAsync(out, d, sizeof(double) * m->ncells, hipMemcpyDeviceToHost, ctx->stream));
    GRMP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, d);
    return GRMP_OK;
}
