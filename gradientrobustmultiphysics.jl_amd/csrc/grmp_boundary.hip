// Inhomogeneous Dirichlet data of libgrmp_hip.so: the boundary values of an
// H1 P1/P2 or H1BR{2} space for polynomial data g, by interpolation or by
// the L2 best approximation on the boundary faces, and the penalized
// right-hand side that carries them into the solve.  At the end: the node
// values of a solution for export.
//
// boundarydata! (src/boundarydata.jl:198-395):
//  * InterpolateDirichletBoundary: the dofs of the boundary faces take the
//    values of g at their points (vertices, P2 face midpoints).
//  * BestapproxDirichletBoundary (:259-352): the mass matrix of the boundary
//    faces (ON_BFACES, Identity x Identity) and the load of g on them are
//    assembled, and the system compressed to the boundary dofs is solved
//    (the reference factorizes it; here a Jacobi-PCG runs to a relative
//    residual of rtol inside one workgroup).
//  * solve_direct! then fixes b_d = penalty * Target_d (src/solvers.jl:647-652).
//  * H1BR{2} (src/fedefs/h1v_br.jl:104-113,186-194): a boundary face carries
//    five dofs, the P1 node dofs of both components ("N1i1") and the face
//    bubble 6 s (1 - s) times the face normal, so the components couple; a
//    face is one unit of the loops below.  Its interpolation keeps the data's
//    normal flux through the face (:37-81): bubble = sum_c (int_F g_c -
//    int_F I g_c) n_c / |F|.
//
// Boundary faces are walked in cell order and local face order (0: nodes
// 0-1, 1: 1-2, 2: 2-0), each face oriented as its cell's local face; a
// boundary dof's row sums its faces in that order, so results do not depend
// on scheduling.  Segment quadrature is Gauss-Legendre with
// ceil((quadorder + 1) / 2) points, at least the 2 (P1) / 3 (P2) that
// integrate the face mass matrix exactly.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "grmp_internal.h"
#include "grmp_square.h"

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#define BD_NCOEF 10  // g(x, y) = sum of c_k m_k, m = 1, x, y, x^2, xy, y^2, x^3, x^2 y, x y^2, y^3
#define BD_MAXCOMP 2 // components of vector-valued data (one polynomial each)
#define BD_MAXQ 4    // Gauss-Legendre points on a segment (exact to degree 7)
#define BD_CG_T 1024

struct BdData {
    double c[BD_MAXCOMP][BD_NCOEF];
};

struct BdMesh {
    const double *coords;     // [2][nnodes]
    const int32_t *cells;     // [3][ncells]
    const int32_t *cellfaces; // [3][ncells] or nullptr (generated unit square)
    const uint8_t *bface;     // [nfaces] or nullptr
    const int32_t *celldofs;  // [nloc][ncells]
    const int8_t *fsign;      // [3][ncells] CellFaceSigns (H1BR{2})
    int64_t nnodes, ncells;
    int N, fe;
    int ncomp, per;           // components of a face unit (H1BR{2}: 1 unit of both); local dofs per component
};

__device__ __forceinline__ double bd_g(const BdData &g, int comp, double x, double y)
{
    const double *c = g.c[comp];
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

// Gauss-Legendre on [0, 1]: points s_q and weights w_q (sum 1)
__device__ __forceinline__ void bd_gauss(int nq, int q, double &s, double &w)
{
    if (nq == 1) {
        s = 0.5;
        w = 1.0;
    } else if (nq == 2) {
        const double d = 0.28867513459481288225;  // 1 / (2 sqrt 3)
        s = q == 0 ? 0.5 - d : 0.5 + d;
        w = 0.5;
    } else if (nq == 3) {
        const double d = 0.38729833462074168852;  // sqrt(3/5) / 2
        s = q == 0 ? 0.5 - d : (q == 1 ? 0.5 : 0.5 + d);
        w = q == 1 ? 4.0 / 9.0 : 5.0 / 18.0;
    } else {
        const double d1 = 0.16999052179242813185, d2 = 0.43056815579702628810;  // roots / 2
        const double w1 = 0.32607257743127307131, w2 = 0.17392742256872692869;  // weights / 2
        s = q == 0 ? 0.5 - d2 : (q == 1 ? 0.5 - d1 : (q == 2 ? 0.5 + d1 : 0.5 + d2));
        w = (q == 0 || q == 3) ? w2 : w1;
    }
}

// 1D trace basis of the face (local dof 0 at the first node, 1 at the second, 2 the P2 midpoint)
__device__ __forceinline__ double bd_phi(int fe, int a, double s)
{
    if (fe == GRMP_FE_H1P1) return a == 0 ? 1.0 - s : s;
    if (a == 0) return (1.0 - s) * (1.0 - 2.0 * s);
    if (a == 1) return s * (2.0 * s - 1.0);
    return 4.0 * s * (1.0 - s);
}

__device__ __forceinline__ bool bd_is_boundary(const BdMesh &M, int64_t c, int k)
{
    if (M.cellfaces) return M.bface[M.cellfaces[k * M.ncells + c]] != 0;
    const int64_t a = M.cells[k * M.ncells + c], b = M.cells[((k + 1) % 3) * M.ncells + c];
    const int64_t W = M.N + 1;
    const int64_t ia = a % W, ja = a / W, ib = b % W, jb = b / W;
    return (ia == 0 && ib == 0) || (ia == M.N && ib == M.N) || (ja == 0 && jb == 0) || (ja == M.N && jb == M.N);
}

#define BD_MAXD 5  // dofs of a face unit (H1BR{2}: 2 nodes x 2 components + the bubble)

// local dofs of component comp on boundary face (c, k): global dof numbers
__device__ __forceinline__ int bd_face_dofs(const BdMesh &M, int64_t c, int k, int comp, int32_t d[BD_MAXD])
{
    const int64_t nc = M.ncells, base = (int64_t)comp * M.per;
    if (M.fe == GRMP_FE_H1BR) {
        d[0] = M.celldofs[k * nc + c];
        d[1] = M.celldofs[((k + 1) % 3) * nc + c];
        d[2] = M.celldofs[(3 + k) * nc + c];
        d[3] = M.celldofs[(3 + (k + 1) % 3) * nc + c];
        d[4] = M.celldofs[(6 + k) * nc + c];
        return 5;
    }
    d[0] = M.celldofs[(base + k) * nc + c];
    d[1] = M.celldofs[(base + (k + 1) % 3) * nc + c];
    if (M.fe == GRMP_FE_H1P2) {
        d[2] = M.celldofs[(base + 3 + k) * nc + c];
        return 3;
    }
    return 2;
}

// a boundary face of one component: v = i ncomp + comp over the face list
__device__ __forceinline__ int bd_vface(const BdMesh &M, const int32_t *bf, int64_t v, int64_t &c, int &k,
                                        int32_t d[BD_MAXD], int &comp)
{
    const int64_t i = v / M.ncomp;
    comp = (int)(v - i * M.ncomp);
    c = bf[i] / 3;
    k = bf[i] - 3 * (int)c;
    return bd_face_dofs(M, c, k, comp, d);
}

__device__ __forceinline__ void bd_face_geom(const BdMesh &M, int64_t c, int k, double &x0, double &y0, double &x1,
                                             double &y1)
{
    const int64_t nc = M.ncells, nn = M.nnodes;
    const int32_t a = M.cells[k * nc + c], b = M.cells[((k + 1) % 3) * nc + c];
    x0 = M.coords[a];
    y0 = M.coords[nn + a];
    x1 = M.coords[b];
    y1 = M.coords[nn + b];
}

// FaceNormals of boundary face (c, k) in its global orientation (CellFaceSigns)
__device__ __forceinline__ void bd_face_normal(const BdMesh &M, int64_t c, int k, double n[2])
{
    const int64_t nc = M.ncells, nn = M.nnodes;
    const int32_t a0 = M.cells[k * nc + c], b0 = M.cells[((k + 1) % 3) * nc + c];
    const bool own = M.fsign[k * nc + c] > 0;
    const int32_t a = own ? a0 : b0, b = own ? b0 : a0;
    const double tx = M.coords[b] - M.coords[a], ty = M.coords[nn + b] - M.coords[nn + a];
    const double len = sqrt(tx * tx + ty * ty);
    n[0] = ty / len;
    n[1] = -tx / len;
}

// H1BR{2} trace basis on the face: P1 per component, then the normal bubble
__device__ __forceinline__ void bd_br_val(int a, double s, const double n[2], double v[2])
{
    const double l = 1.0 - s;
    if (a < 4) {
        const double p = (a & 1) ? s : l;
        v[0] = a < 2 ? p : 0.0;
        v[1] = a < 2 ? 0.0 : p;
    } else {
        const double bub = 6.0 * s * l;
        v[0] = bub * n[0];
        v[1] = bub * n[1];
    }
}

// ------------------------------------------------------------ boundary faces
__global__ void k_bd_flag(BdMesh M, int32_t *__restrict__ flag)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * M.ncells) return;
    const int64_t c = t / 3;
    const int k = (int)(t - 3 * c);
    flag[t] = bd_is_boundary(M, c, k) ? 1 : 0;
}

__global__ void k_bd_compact(const int32_t *__restrict__ flag, const int32_t *__restrict__ pos, int64_t n3,
                             int32_t *__restrict__ bf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n3 && flag[t]) bf[pos[t]] = (int32_t)t;  // t = 3 c + k
}

// boundary dof flags (int32 for the scan)
__global__ void k_bd_dofflag(BdMesh M, const int32_t *__restrict__ bf, int64_t nv, int32_t *__restrict__ dflag)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int64_t c;
    int k, comp;
    int32_t d[BD_MAXD];
    const int nd = bd_vface(M, bf, v, c, k, d, comp);
    for (int a = 0; a < nd; ++a) dflag[d[a]] = 1;
}

__global__ void k_bd_rows(const int32_t *__restrict__ dflag, const int32_t *__restrict__ dpos, int64_t ndofs,
                          int32_t *__restrict__ comp, int32_t *__restrict__ bdof)
{
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndofs) return;
    if (dflag[d]) {
        comp[d] = dpos[d];
        bdof[dpos[d]] = (int32_t)d;
    } else {
        comp[d] = -1;
    }
}

// incidences (boundary face, local dof) per compressed row: count, fill, sort
__global__ void k_bd_count(BdMesh M, const int32_t *__restrict__ bf, int64_t nv, const int32_t *__restrict__ comp,
                           int32_t *__restrict__ cnt)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int64_t c;
    int k, cp;
    int32_t d[BD_MAXD];
    const int nd = bd_vface(M, bf, v, c, k, d, cp);
    for (int a = 0; a < nd; ++a) atomicAdd(&cnt[comp[d[a]]], 1);
}

__global__ void k_bd_fill(BdMesh M, const int32_t *__restrict__ bf, int64_t nv, const int32_t *__restrict__ comp,
                          const int32_t *__restrict__ ptr, int32_t *__restrict__ cur, int32_t *__restrict__ inc)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int64_t c;
    int k, cp;
    int32_t d[BD_MAXD];
    const int nd = bd_vface(M, bf, v, c, k, d, cp);
    for (int a = 0; a < nd; ++a) {
        const int32_t r = comp[d[a]];
        const int32_t slot = ptr[r] + atomicAdd(&cur[r], 1);
        inc[slot] = (int32_t)(v * 8 + a);  // component face index, local dof
    }
}

__global__ void k_bd_sort(const int32_t *__restrict__ ptr, int64_t nb, int32_t *__restrict__ inc)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nb) return;
    for (int32_t p = ptr[r] + 1; p < ptr[r + 1]; ++p) {
        const int32_t v = inc[p];
        int32_t q = p - 1;
        while (q >= ptr[r] && inc[q] > v) {
            inc[q + 1] = inc[q];
            --q;
        }
        inc[q + 1] = v;
    }
}

// ------------------------------------------------------ best approximation
struct BdSys {
    BdMesh M;
    BdData g;
    const int32_t *bf, *ptr, *inc, *comp;
    int64_t nb;
    int nq;
};

__device__ __forceinline__ void bd_face_len(const BdSys &S, int64_t i, double &len, double &x0, double &y0, double &x1,
                                            double &y1)
{
    const int64_t c = S.bf[i] / 3;
    const int k = S.bf[i] - 3 * (int)c;
    bd_face_geom(S.M, c, k, x0, y0, x1, y1);
    const double dx = x1 - x0, dy = y1 - y0;
    len = sqrt(dx * dx + dy * dy);
}

// (M x)_r and the diagonal M_rr, row r summed over its faces in order
__device__ __forceinline__ double bd_row_apply(const BdSys &S, int64_t r, const double *x, double *diag)
{
    double y = 0.0, dg = 0.0;
    for (int32_t p = S.ptr[r]; p < S.ptr[r + 1]; ++p) {
        const int64_t v = S.inc[p] >> 3;
        const int a = S.inc[p] & 7;
        double len, x0, y0, x1, y1;
        bd_face_len(S, v / S.M.ncomp, len, x0, y0, x1, y1);
        int64_t c;
        int k, comp;
        int32_t d[BD_MAXD];
        const int nd = bd_vface(S.M, S.bf, v, c, k, d, comp);
        double n[2] = {0.0, 0.0};
        if (S.M.fe == GRMP_FE_H1BR) bd_face_normal(S.M, c, k, n);
        for (int b = 0; b < nd; ++b) {
            double m = 0.0;
            for (int q = 0; q < S.nq; ++q) {
                double s, w;
                bd_gauss(S.nq, q, s, w);
                if (S.M.fe == GRMP_FE_H1BR) {
                    double va[2], vb[2];
                    bd_br_val(a, s, n, va);
                    bd_br_val(b, s, n, vb);
                    double t = 0.0;
                    t += va[0] * vb[0];
                    t += va[1] * vb[1];
                    m += w * t;
                } else {
                    m += w * (bd_phi(S.M.fe, a, s) * bd_phi(S.M.fe, b, s));
                }
            }
            m *= len;
            if (x) y += m * x[S.comp[d[b]]];
            if (b == a) dg += m;
        }
    }
    if (diag) *diag = dg;
    return y;
}

__device__ __forceinline__ double bd_row_load(const BdSys &S, int64_t r)
{
    double f = 0.0;
    for (int32_t p = S.ptr[r]; p < S.ptr[r + 1]; ++p) {
        const int64_t v = S.inc[p] >> 3;
        const int a = S.inc[p] & 7;
        const int comp = (int)(v % S.M.ncomp);
        double len, x0, y0, x1, y1;
        bd_face_len(S, v / S.M.ncomp, len, x0, y0, x1, y1);
        double n[2] = {0.0, 0.0};
        if (S.M.fe == GRMP_FE_H1BR) {
            const int64_t i = v / S.M.ncomp, c = S.bf[i] / 3;
            bd_face_normal(S.M, c, S.bf[i] - 3 * (int)c, n);
        }
        double l = 0.0;
        for (int q = 0; q < S.nq; ++q) {
            double s, w;
            bd_gauss(S.nq, q, s, w);
            const double xq = x0 + s * (x1 - x0), yq = y0 + s * (y1 - y0);
            if (S.M.fe == GRMP_FE_H1BR) {
                double va[2];
                bd_br_val(a, s, n, va);
                double t = 0.0;
                t += bd_g(S.g, 0, xq, yq) * va[0];
                t += bd_g(S.g, 1, xq, yq) * va[1];
                l += w * t;
            } else {
                l += w * (bd_g(S.g, comp, xq, yq) * bd_phi(S.M.fe, a, s));
            }
        }
        f += l * len;
    }
    return f;
}

__device__ __forceinline__ double bd_block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < BD_CG_T / 64; ++k) s += red[k];
    return s;
}

// Jacobi-PCG on the compressed boundary system in one workgroup (the loop
// condition is uniform: every thread holds the same block sums); x stays in
// xv, status {iterations, relres} goes to st.
__global__ __launch_bounds__(BD_CG_T) void k_bd_cg(BdSys S, double rtol, int maxit, double *__restrict__ xv,
                                                   double *__restrict__ rv, double *__restrict__ zv,
                                                   double *__restrict__ pv, double *__restrict__ dv,
                                                   double *__restrict__ st)
{
    __shared__ double red[BD_CG_T / 64];
    const int64_t nb = S.nb;
    double rz = 0.0, bb = 0.0;
    for (int64_t r = threadIdx.x; r < nb; r += BD_CG_T) {
        double dg;
        bd_row_apply(S, r, nullptr, &dg);
        const double b = bd_row_load(S, r);
        dv[r] = 1.0 / dg;
        xv[r] = 0.0;
        rv[r] = b;
        zv[r] = b / dg;
        pv[r] = b / dg;
        rz += b * (b / dg);
        bb += b * b;
    }
    rz = bd_block_sum(rz, red);
    bb = bd_block_sum(bb, red);
    double rr = bb;
    int it = 0;
    while (it < maxit && rr > rtol * rtol * bb && rr > 0.0) {
        __syncthreads();  // p complete before the products
        double pq = 0.0;
        for (int64_t r = threadIdx.x; r < nb; r += BD_CG_T) {
            const double q = bd_row_apply(S, r, pv, nullptr);
            zv[r] = q;  // q kept in z until the update below
            pq += pv[r] * q;
        }
        pq = bd_block_sum(pq, red);
        const double alpha = rz / pq;
        double rz2 = 0.0, rr2 = 0.0;
        for (int64_t r = threadIdx.x; r < nb; r += BD_CG_T) {
            xv[r] += alpha * pv[r];
            const double ri = rv[r] - alpha * zv[r];
            rv[r] = ri;
            const double zi = dv[r] * ri;
            zv[r] = zi;
            rz2 += ri * zi;
            rr2 += ri * ri;
        }
        rz2 = bd_block_sum(rz2, red);
        rr = bd_block_sum(rr2, red);
        const double beta = rz2 / rz;
        rz = rz2;
        for (int64_t r = threadIdx.x; r < nb; r += BD_CG_T) pv[r] = zv[r] + beta * pv[r];
        ++it;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        st[0] = (double)it;
        st[1] = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    }
}

__global__ void k_bd_scatter(const double *__restrict__ xv, const int32_t *__restrict__ bdof, int64_t nb,
                             double *__restrict__ target, int64_t offset)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nb) target[offset + bdof[r]] = xv[r];
}

// interpolation: g at the dof points of the boundary faces
__global__ void k_bd_interp(BdMesh M, BdData g, const int32_t *__restrict__ bf, int64_t nv,
                            double *__restrict__ target, int64_t offset)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int64_t c;
    int k, comp;
    int32_t d[BD_MAXD];
    const int nd = bd_vface(M, bf, v, c, k, d, comp);
    double x0, y0, x1, y1;
    bd_face_geom(M, c, k, x0, y0, x1, y1);
    if (M.fe == GRMP_FE_H1BR) {
        // nodes, then the bubble that keeps the normal flux of g through the face
        const double g0[2] = {bd_g(g, 0, x0, y0), bd_g(g, 1, x0, y0)};
        const double g1[2] = {bd_g(g, 0, x1, y1), bd_g(g, 1, x1, y1)};
        target[offset + d[0]] = g0[0];
        target[offset + d[1]] = g1[0];
        target[offset + d[2]] = g0[1];
        target[offset + d[3]] = g1[1];
        double n[2];
        bd_face_normal(M, c, k, n);
        const double dx = x1 - x0, dy = y1 - y0, len = sqrt(dx * dx + dy * dy);
        double val = 0.0;
        for (int cc = 0; cc < 2; ++cc) {
            double mean = 0.0;  // int_F g_c: 4-point Gauss (exact to degree 7)
            for (int q = 0; q < BD_MAXQ; ++q) {
                double s, w;
                bd_gauss(BD_MAXQ, q, s, w);
                mean += w * bd_g(g, cc, x0 + s * dx, y0 + s * dy);
            }
            mean *= len;
            double p1 = 0.0;
            p1 += g0[cc] * len / 2.0;
            p1 += g1[cc] * len / 2.0;
            val += (mean - p1) * n[cc];
        }
        target[offset + d[4]] = val / len;
        return;
    }
    // a dof shared by two faces gets the same value from both (its point)
    target[offset + d[0]] = bd_g(g, comp, x0, y0);
    target[offset + d[1]] = bd_g(g, comp, x1, y1);
    if (nd == 3) target[offset + d[2]] = bd_g(g, comp, 0.5 * (x0 + x1), 0.5 * (y0 + y1));
}

__global__ void k_penalize_rhs(const uint8_t *__restrict__ fixed, double penalty, const double *__restrict__ target,
                               double *__restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && fixed[i]) b[i] = penalty * target[i];
}

// ------------------------------------------------------------------ host
struct BdFaces {
    int32_t *bf = nullptr;
    int64_t nbf = 0;
};

static int bd_check(grmp_ctx *ctx, const grmp_space *s, const double *coef, int ncoef, const grmp_vector *target,
                    int64_t offset)
{
    GRMP_CHECK(ctx, s->mesh->dim == 2 && s->mesh->nloc == 3, GRMP_ERR_UNSUPPORTED,
               "Dirichlet data: triangle meshes only");
    GRMP_CHECK(ctx, ((s->fe == GRMP_FE_H1P1 || s->fe == GRMP_FE_H1P2) && s->ncomp >= 1 && s->ncomp <= BD_MAXCOMP) ||
                        (s->fe == GRMP_FE_H1BR && s->ncomp == 2 && s->mesh->fsign),
               GRMP_ERR_UNSUPPORTED, "Dirichlet data: H1P1 / H1P2 spaces with 1 or 2 components and H1BR{2} only");
    GRMP_CHECK(ctx, s->mesh->cellfaces || s->mesh->kind == MESH_UNITSQUARE, GRMP_ERR_UNSUPPORTED,
               "Dirichlet data: the mesh has no face adjacency");
    GRMP_CHECK(ctx, coef && ncoef >= 1 && ncoef <= BD_NCOEF, GRMP_ERR_ARG,
               "Dirichlet data: 1..%d polynomial coefficients", BD_NCOEF);
    GRMP_CHECK(ctx, target && target->ctx == ctx && offset >= 0 && offset + s->ndofs <= target->n, GRMP_ERR_ARG,
               "Dirichlet data: target too short for the space at offset %lld", (long long)offset);
    return GRMP_OK;
}

static BdMesh bd_mesh(const grmp_space *s)
{
    const grmp_mesh *m = s->mesh;
    BdMesh M;
    M.coords = m->coords;
    M.cells = m->cells;
    M.cellfaces = m->cellfaces;
    M.bface = m->bface;
    M.celldofs = s->celldofs;
    M.fsign = m->fsign;
    M.nnodes = m->nnodes;
    M.ncells = m->ncells;
    M.N = m->N;
    M.fe = s->fe;
    M.ncomp = s->fe == GRMP_FE_H1BR ? 1 : s->ncomp;
    M.per = s->fe == GRMP_FE_H1P2 ? 6 : 3;
    return M;
}

static int bd_faces(grmp_ctx *ctx, const BdMesh &M, BdFaces &F)
{
    const int64_t n3 = 3 * M.ncells;
    int32_t *flag = nullptr, *pos = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &flag, n3)) || (rc = dalloc(ctx, &pos, n3 + 1))) break;
        hipLaunchKernelGGL(k_bd_flag, dim3(nblk(n3, 256)), dim3(256), 0, ctx->stream, M, flag);
        if ((rc = exclusive_scan_i32(ctx, flag, pos, n3))) break;
        int32_t nbf = 0;
        if (hipMemcpyAsync(&nbf, pos + n3, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "Dirichlet data: boundary face count");
            break;
        }
        F.nbf = nbf;
        if ((rc = dalloc(ctx, &F.bf, std::max<int64_t>(nbf, 1)))) break;
        hipLaunchKernelGGL(k_bd_compact, dim3(nblk(n3, 256)), dim3(256), 0, ctx->stream, flag, pos, n3, F.bf);
        if (hipGetLastError() != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "Dirichlet data: compaction");
    } while (0);
    dfree(ctx, flag);
    dfree(ctx, pos);
    return rc;
}

extern "C" int grmp_boundary_interpolate(const grmp_space *s, const double *coef, int ncoef, grmp_vector *target,
                                         int64_t offset)
{
    if (!s) return GRMP_ERR_ARG;
    grmp_ctx *ctx = s->ctx;
    GRMP_TRY(bd_check(ctx, s, coef, ncoef, target, offset));
    std::lock_guard<std::mutex> g(ctx->mu);
    BdData gd;
    memset(&gd, 0, sizeof(gd));
    for (int m = 0; m < s->ncomp; ++m)
        for (int k = 0; k < ncoef; ++k) gd.c[m][k] = coef[m * ncoef + k];
    const BdMesh M = bd_mesh(s);
    BdFaces F;
    int rc = bd_faces(ctx, M, F);
    const int64_t nv = F.nbf * M.ncomp;
    if (!rc && nv) {
        hipLaunchKernelGGL(k_bd_interp, dim3(nblk(nv, 256)), dim3(256), 0, ctx->stream, M, gd, F.bf, nv, target->d,
                           offset);
        if (hipGetLastError() != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "Dirichlet data: interpolation");
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = grmp_fail(ctx, GRMP_ERR_HIP, "Dirichlet data");
    dfree(ctx, F.bf);
    return rc;
}

extern "C" int grmp_boundary_bestapprox(const grmp_space *s, const double *coef, int ncoef, int quadorder,
                                        double rtol, grmp_vector *target, int64_t offset, int64_t *iterations,
                                        double *relres)
{
    if (!s) return GRMP_ERR_ARG;
    grmp_ctx *ctx = s->ctx;
    GRMP_TRY(bd_check(ctx, s, coef, ncoef, target, offset));
    GRMP_CHECK(ctx, quadorder >= 0 && quadorder <= 2 * BD_MAXQ - 1, GRMP_ERR_UNSUPPORTED,
               "Dirichlet data: segment quadrature order %d above %d", quadorder, 2 * BD_MAXQ - 1);
    GRMP_CHECK(ctx, rtol > 0.0 && rtol < 1.0, GRMP_ERR_ARG, "Dirichlet data: rtol must lie in (0, 1)");
    std::lock_guard<std::mutex> g(ctx->mu);
    BdSys S;
    memset(&S, 0, sizeof(S));
    for (int m = 0; m < s->ncomp; ++m)
        for (int k = 0; k < ncoef; ++k) S.g.c[m][k] = coef[m * ncoef + k];
    S.M = bd_mesh(s);
    // the mass matrix needs degree 2 p exactly, the load p + the data's degree
    S.nq = std::max((quadorder + 2) / 2, s->fe == GRMP_FE_H1P1 ? 2 : 3);
    const int64_t nd = s->ndofs;
    BdFaces F;
    int32_t *dflag = nullptr, *dpos = nullptr, *comp = nullptr, *bdof = nullptr, *cnt = nullptr, *ptr = nullptr,
            *cur = nullptr, *inc = nullptr;
    double *work = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = bd_faces(ctx, S.M, F))) break;
        if (F.nbf == 0) break;
        const int per = s->fe == GRMP_FE_H1BR ? 5 : (s->fe == GRMP_FE_H1P2 ? 3 : 2);
        const int64_t nv = F.nbf * S.M.ncomp;  // boundary face units
        if ((rc = dalloc(ctx, &dflag, nd)) || (rc = dalloc(ctx, &dpos, nd + 1)) || (rc = dalloc(ctx, &comp, nd)))
            break;
        GRMP_HIP(ctx, hipMemsetAsync(dflag, 0, sizeof(int32_t) * nd, ctx->stream));
        hipLaunchKernelGGL(k_bd_dofflag, dim3(nblk(nv, 256)), dim3(256), 0, ctx->stream, S.M, F.bf, nv, dflag);
        if ((rc = exclusive_scan_i32(ctx, dflag, dpos, nd))) break;
        int32_t nb = 0;
        if (hipMemcpyAsync(&nb, dpos + nd, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "Dirichlet data: boundary dof count");
            break;
        }
        S.nb = nb;
        const int64_t ninc = nv * per;
        if ((rc = dalloc(ctx, &bdof, nb)) || (rc = dalloc(ctx, &cnt, nb)) || (rc = dalloc(ctx, &ptr, nb + 1)) ||
            (rc = dalloc(ctx, &cur, nb)) || (rc = dalloc(ctx, &inc, ninc)) || (rc = dalloc(ctx, &work, 5 * (size_t)nb + 2)))
            break;
        hipLaunchKernelGGL(k_bd_rows, dim3(nblk(nd, 256)), dim3(256), 0, ctx->stream, dflag, dpos, nd, comp, bdof);
        GRMP_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int32_t) * nb, ctx->stream));
        GRMP_HIP(ctx, hipMemsetAsync(cur, 0, sizeof(int32_t) * nb, ctx->stream));
        hipLaunchKernelGGL(k_bd_count, dim3(nblk(nv, 256)), dim3(256), 0, ctx->stream, S.M, F.bf, nv, comp, cnt);
        if ((rc = exclusive_scan_i32(ctx, cnt, ptr, nb))) break;
        hipLaunchKernelGGL(k_bd_fill, dim3(nblk(nv, 256)), dim3(256), 0, ctx->stream, S.M, F.bf, nv, comp, ptr,
                           cur, inc);
        hipLaunchKernelGGL(k_bd_sort, dim3(nblk(nb, 256)), dim3(256), 0, ctx->stream, ptr, (int64_t)nb, inc);
        S.bf = F.bf;
        S.ptr = ptr;
        S.inc = inc;
        S.comp = comp;
        double *xv = work, *rv = work + nb, *zv = work + 2 * (size_t)nb, *pv = work + 3 * (size_t)nb,
               *dv = work + 4 * (size_t)nb, *st = work + 5 * (size_t)nb;
        const int maxit = (int)std::min<int64_t>(100000, 10 * (int64_t)nb + 100);
        hipLaunchKernelGGL(k_bd_cg, dim3(1), dim3(BD_CG_T), 0, ctx->stream, S, rtol, maxit, xv, rv, zv, pv, dv, st);
        hipLaunchKernelGGL(k_bd_scatter, dim3(nblk(nb, 256)), dim3(256), 0, ctx->stream, xv, bdof, (int64_t)nb,
                           target->d, offset);
        GRMP_HIP(ctx, hipGetLastError());
        double hst[2] = {0.0, 0.0};
        if (hipMemcpyAsync(hst, st, sizeof(hst), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "Dirichlet data: boundary solve");
            break;
        }
        if (iterations) *iterations = (int64_t)hst[0];
        if (relres) *relres = hst[1];
        if (!(hst[1] <= rtol)) rc = grmp_fail(ctx, GRMP_ERR_NOCONV, "Dirichlet data: boundary solve reached relres %.3e",
                                              hst[1]);
    } while (0);
    dfree(ctx, F.bf);
    dfree(ctx, dflag);
    dfree(ctx, dpos);
    dfree(ctx, comp);
    dfree(ctx, bdof);
    dfree(ctx, cnt);
    dfree(ctx, ptr);
    dfree(ctx, cur);
    dfree(ctx, inc);
    dfree(ctx, work);
    return rc;
}

extern "C" int grmp_matrix_penalize_rhs(const grmp_matrix *A, const grmp_vector *target, grmp_vector *b)
{
    if (!A || !target || !b) return GRMP_ERR_ARG;
    grmp_ctx *ctx = A->ctx;
    GRMP_CHECK(ctx, A->fixed && target->n == A->n && b->n == A->n, GRMP_ERR_ARG,
               "grmp_matrix_penalize_rhs: fixed dofs and vectors of the system's length %lld needed", (long long)A->n);
    std::lock_guard<std::mutex> g(ctx->mu);
    hipLaunchKernelGGL(k_penalize_rhs, dim3(nblk(A->n, 256)), dim3(256), 0, ctx->stream, A->fixed, A->penalty,
                       target->d, b->d, A->n);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}

// ------------------------------------------------------------ node values
// nodevalues!(target, source, FE, Identity) with continuous = false
// (src/interpolations.jl:704-812; L2P0: src/fedefs/l2_p0.jl:52-68): per node
// and component the values of the cells around the node at that vertex,
// summed in cell order and divided by their number.  At a vertex every
// basis function of P1 / P2 / BR is exactly 0 or 1 (the BR bubbles vanish),
// so a cell's value there is the coefficient of the vertex's dof; for P0 it
// is the cell's coefficient.
__global__ void k_node_count(const int32_t *__restrict__ cells, int64_t nc, int32_t *__restrict__ cnt)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * nc) atomicAdd(&cnt[cells[t]], 1);
}

__global__ void k_node_fill(const int32_t *__restrict__ cells, int64_t nc, const int32_t *__restrict__ ptr,
                            int32_t *__restrict__ cur, int32_t *__restrict__ nc_inc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nc) return;
    const int32_t n = cells[t];
    const int64_t c = t % nc, k = t / nc;  // cells is SoA [3][nc]
    nc_inc[ptr[n] + atomicAdd(&cur[n], 1)] = (int32_t)(c * 4 + k);
}

__global__ void k_nodevalues(const int32_t *__restrict__ ptr, const int32_t *__restrict__ nc_inc,
                             const int32_t *__restrict__ celldofs, int64_t nc, int64_t nnodes, int fe, int ncomp,
                             int per_comp, const double *__restrict__ u, double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnodes * ncomp) return;
    const int64_t n = t / ncomp;
    const int c = (int)(t - n * ncomp);
    double acc = 0.0;
    for (int32_t p = ptr[n]; p < ptr[n + 1]; ++p) {
        const int64_t cell = nc_inc[p] >> 2;
        const int k = nc_inc[p] & 3;
        const int loc = fe == GRMP_FE_L2P0 ? c : c * per_comp + k;
        acc += u[celldofs[(int64_t)loc * nc + cell]];
    }
    const int32_t cnt = ptr[n + 1] - ptr[n];
    out[t] = cnt > 0 ? acc / (double)cnt : 0.0;
}

extern "C" int grmp_nodevalues(const grmp_space *s, const grmp_vector *u, int64_t offset, double *out)
{
    if (!s || !u || !out) return GRMP_ERR_ARG;
    grmp_ctx *ctx = s->ctx;
    const grmp_mesh *m = s->mesh;
    GRMP_CHECK(ctx, m->dim == 2 && m->nloc == 3, GRMP_ERR_UNSUPPORTED, "grmp_nodevalues: triangle meshes only");
    GRMP_CHECK(ctx, s->fe == GRMP_FE_H1P1 || s->fe == GRMP_FE_H1P2 || s->fe == GRMP_FE_H1BR || s->fe == GRMP_FE_L2P0,
               GRMP_ERR_UNSUPPORTED, "grmp_nodevalues: FE %d is not built", s->fe);
    GRMP_CHECK(ctx, u->ctx == ctx && offset >= 0 && offset + s->ndofs <= u->n, GRMP_ERR_ARG,
               "grmp_nodevalues: coefficient vector too short at offset %lld", (long long)offset);
    const int ncomp = s->fe == GRMP_FE_H1BR ? 2 : s->ncomp;
    const int per_comp = s->fe == GRMP_FE_H1P2 ? 6 : 3;
    std::lock_guard<std::mutex> g(ctx->mu);
    const int64_t nc = m->ncells, nn = m->nnodes;
    int32_t *cnt = nullptr, *ptr = nullptr, *cur = nullptr, *inc = nullptr;
    double *dout = nullptr;
    int rc = GRMP_OK;
    do {
        if ((rc = dalloc(ctx, &cnt, nn)) || (rc = dalloc(ctx, &ptr, nn + 1)) || (rc = dalloc(ctx, &cur, nn)) ||
            (rc = dalloc(ctx, &inc, 3 * nc)) || (rc = dalloc(ctx, &dout, nn * ncomp)))
            break;
        GRMP_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int32_t) * nn, ctx->stream));
        GRMP_HIP(ctx, hipMemsetAsync(cur, 0, sizeof(int32_t) * nn, ctx->stream));
        hipLaunchKernelGGL(k_node_count, dim3(nblk(3 * nc, 256)), dim3(256), 0, ctx->stream, m->cells, nc, cnt);
        if ((rc = exclusive_scan_i32(ctx, cnt, ptr, nn))) break;
        hipLaunchKernelGGL(k_node_fill, dim3(nblk(3 * nc, 256)), dim3(256), 0, ctx->stream, m->cells, nc, ptr, cur, inc);
        hipLaunchKernelGGL(k_bd_sort, dim3(nblk(nn, 256)), dim3(256), 0, ctx->stream, ptr, nn, inc);
        hipLaunchKernelGGL(k_nodevalues, dim3(nblk(nn * ncomp, 256)), dim3(256), 0, ctx->stream, ptr, inc,
                           s->celldofs, nc, nn, s->fe, ncomp, per_comp, u->d + offset, dout);
        GRMP_HIP(ctx, hipGetLastError());
        if (hipMemcpyAsync(out, dout, sizeof(double) * nn * ncomp, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = grmp_fail(ctx, GRMP_ERR_HIP, "grmp_nodevalues: download");
    } while (0);
    dfree(ctx, cnt);
    dfree(ctx, ptr);
    dfree(ctx, cur);
    dfree(ctx, inc);
    dfree(ctx, dout);
    return rc;
}

// ------------------------------------------------------------------ axpy
__global__ void k_axpy(double a, const double *__restrict__ x, double *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

extern "C" int grmp_vector_axpy(double a, const grmp_vector *x, grmp_vector *y)
{
    if (!x || !y) return GRMP_ERR_ARG;
    grmp_ctx *ctx = y->ctx;
    GRMP_CHECK(ctx, x->ctx == ctx && x->n == y->n, GRMP_ERR_ARG, "grmp_vector_axpy: lengths %lld and %lld differ",
               (long long)x->n, (long long)y->n);
    std::lock_guard<std::mutex> g(ctx->mu);
    if (y->n == 0) return GRMP_OK;
    hipLaunchKernelGGL(k_axpy, dim3(nblk(y->n, 256)), dim3(256), 0, ctx->stream, a, x->d, y->d, y->n);
    GRMP_HIP(ctx, hipGetLastError());
    return GRMP_OK;
}
