/*
 * grmp_oracle.c -- CPU restatement of GradientRobustMultiPhysics.jl's assembly
 * and solve hot path.  TEST INFRASTRUCTURE ONLY: this file is the checker that
 * the HIP library is compared against (tests/, the smoke check and the CPU
 * baseline leg of the benchmark).  Nothing in the product links or calls it.
 *
 * Compiled with -ffp-contract=off: Julia never fuses a*b+c, so every product
 * and sum below is rounded separately, in the order the reference writes it.
 *
 * What follows which reference file (paths relative to the reference root):
 *   quadrature points/weights ...... src/quadrature.jl:173-195 (Triangle2D),
 *                                    src/quadrature.jl:268-281 (Tetrahedron3D)
 *   reference bases ................ src/fedefs/h1_p1.jl:64-75,
 *                                    src/fedefs/h1_p2.jl:134-145
 *   derivative tables .............. src/feevaluator.jl:235-293 (closed form here)
 *   affine map + A^{-T} ............ src/feevaluator.jl:371-380 (ExtendableGrids
 *                                    update_trafo!/mapderiv!, restated)
 *   Gradient / SymmetricGradient ... src/feevaluator_h1.jl:61-74, 97-116
 *   BilinearForm element loop ...... src/assemblypatterns/bilinearform.jl:180-220,
 *                                    226-377
 *   _addnz zero skip + CSC ......... src/fematrix.jl:54-65 (ExtendableSparse
 *                                    rawupdateindex!/flush!, restated)
 *   LinearForm element loop ........ src/assemblypatterns/linearform.jl:130-233
 *   ItemIntegrator ................. src/assemblypatterns/itemintegrator.jl:243-307
 *   Hooke tensor action ............ src/pdeoperators.jl:256-276
 *   penalty Dirichlet .............. src/solvers.jl:647-652, src/fematrix.jl:349-355
 *   face numbering (CellFaces) ..... ExtendableGrids (not in the reference tree),
 *                                    restated as first appearance in cell order
 *   dofmaps "N1", "N1F1" ........... src/dofmaps.jl:201-363
 *
 * Parity pins (what checks each part) are listed in oracle/README and in the
 * CPU tests under tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_FE_P1 1
#define OR_FE_P2 2
#define OR_FE_BR 3
#define OR_FE_P0 4

#define OR_BLF_LAPLACE 1   /* (kappa grad u, grad v), APT_SymmetricBilinearForm */
#define OR_BLF_HOOKE2D 2   /* (C eps(u), eps(v)), APT_BilinearForm, P1{2} */
#define OR_BLF_MASS 3      /* (kappa u, v), APT_SymmetricBilinearForm */

#define OR_COEF_FACTOR 0   /* NoAction, constant factor (LaplaceOperator(kappa)) */
#define OR_COEF_CELL 1     /* item-dependent Action: result = kappa[cell] .* input */

/* ------------------------------------------------------------------ meshes */

/* Unit square, N x N squares, each split along its (0,0)-(1,1) diagonal into
 * [(i,j),(i+1,j),(i+1,j+1)] and [(i,j),(i+1,j+1),(i,j+1)].  Node id j(N+1)+i,
 * coordinates i/N (correctly rounded).  coords: [nnodes][2], cells: [nc][3]. */
void or_unitsquare(int N, double *coords, int32_t *cells)
{
    int n1 = N + 1;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) {
            coords[2 * (j * n1 + i) + 0] = (double)i / (double)N;
            coords[2 * (j * n1 + i) + 1] = (double)j / (double)N;
        }
    int64_t c = 0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            int32_t v00 = j * n1 + i, v10 = v00 + 1, v01 = v00 + n1, v11 = v01 + 1;
            cells[3 * c + 0] = v00; cells[3 * c + 1] = v10; cells[3 * c + 2] = v11; ++c;
            cells[3 * c + 0] = v00; cells[3 * c + 1] = v11; cells[3 * c + 2] = v01; ++c;
        }
}

/* ------------------------------------------------- face (edge) numbering */

/* Faces of a triangle mesh numbered by first appearance in cell order, local
 * faces [1 2; 2 3; 3 1].  cellfaces: [nc][3]; facenodes: [nf][2] (as first
 * met); facecells: [nf][2] (second = -1 on the boundary).  Returns nfaces. */
int64_t or_tri_faces(int64_t nc, int64_t nnodes, const int32_t *cells,
                     int32_t *cellfaces, int32_t *facenodes, int32_t *facecells)
{
    /* per-node list of (other node, face id) built on the fly */
    int64_t cap = 3 * nc + 16;
    int32_t *head = (int32_t *)malloc(sizeof(int32_t) * nnodes);
    int32_t *next = (int32_t *)malloc(sizeof(int32_t) * cap);
    int32_t *other = (int32_t *)malloc(sizeof(int32_t) * cap);
    int32_t *fid = (int32_t *)malloc(sizeof(int32_t) * cap);
    for (int64_t v = 0; v < nnodes; ++v) head[v] = -1;
    int64_t nf = 0, nl = 0;
    static const int lf[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int64_t c = 0; c < nc; ++c)
        for (int f = 0; f < 3; ++f) {
            int32_t a = cells[3 * c + lf[f][0]], b = cells[3 * c + lf[f][1]];
            int32_t lo = a < b ? a : b, hi = a < b ? b : a;
            int32_t found = -1;
            for (int32_t p = head[lo]; p >= 0; p = next[p])
                if (other[p] == hi) { found = fid[p]; break; }
            if (found < 0) {
                found = (int32_t)nf;
                facenodes[2 * nf + 0] = a; facenodes[2 * nf + 1] = b;
                facecells[2 * nf + 0] = (int32_t)c; facecells[2 * nf + 1] = -1;
                ++nf;
                other[nl] = hi; fid[nl] = found; next[nl] = head[lo]; head[lo] = (int32_t)nl; ++nl;
            } else {
                facecells[2 * found + 1] = (int32_t)c;
            }
            cellfaces[3 * c + f] = found;
        }
    free(head); free(next); free(other); free(fid);
    return nf;
}

/* --------------------------------------------------------------- dofmaps */

/* CellDofs for pattern "N1" (P1, ncomp components, component offset nnodes)
 * and "N1F1" (P2: nodes then faces, face dof = nnodes + face id), following
 * src/dofmaps.jl:201-363 with component blocking (offset4component = coffset).
 * celldofs: [nc][nloc*ncomp]. */
void or_celldofs(int fe, int ncomp, int64_t nc, int64_t nnodes, int64_t nfaces,
                 const int32_t *cells, const int32_t *cellfaces, int32_t *celldofs)
{
    if (fe == OR_FE_P1) {
        int64_t coff = nnodes;
        for (int64_t c = 0; c < nc; ++c)
            for (int k = 0; k < ncomp; ++k)
                for (int a = 0; a < 3; ++a)
                    celldofs[c * 3 * ncomp + k * 3 + a] = (int32_t)(cells[3 * c + a] + k * coff);
    } else if (fe == OR_FE_P2) {
        int64_t coff = nnodes + nfaces;
        for (int64_t c = 0; c < nc; ++c)
            for (int k = 0; k < ncomp; ++k) {
                for (int a = 0; a < 3; ++a)
                    celldofs[c * 6 * ncomp + k * 6 + a] = (int32_t)(cells[3 * c + a] + k * coff);
                for (int f = 0; f < 3; ++f)
                    celldofs[c * 6 * ncomp + k * 6 + 3 + f] = (int32_t)(nnodes + cellfaces[3 * c + f] + k * coff);
            }
    }
}

/* --------------------------------------------------- quadrature + tables */

/* Triangle2D rules (src/quadrature.jl:173-195): order<=1 midpoint, order 2 face
 * midpoints [(1/2,1/2),(0,1/2),(1/2,0)] with weights 1/3.  Weights sum to 1;
 * the physical weight is w*vol. */
int or_quad_tri(int order, double *xref /*[nq][2]*/, double *w)
{
    if (order <= 1) {
        double t = 1.0 / 3.0;
        xref[0] = t; xref[1] = t; w[0] = 1.0;
        return 1;
    }
    if (order == 2) {
        xref[0] = 0.5; xref[1] = 0.5;
        xref[2] = 0.0; xref[3] = 0.5;
        xref[4] = 0.5; xref[5] = 0.0;
        w[0] = w[1] = w[2] = 1.0 / 3.0;
        return 3;
    }
    return 0;
}

/* P1 reference basis on the triangle (src/fedefs/h1_p1.jl:64-75): phi_1 is
 * built as 1 - x1 - x2 by successive subtraction. */
static void p1_tri_vals(const double *x, double *phi)
{
    double l = 1.0;
    l -= x[0]; phi[1] = x[0];
    l -= x[1]; phi[2] = x[1];
    phi[0] = l;
}
static void p1_tri_ders(double d[3][2])
{
    d[0][0] = -1.0; d[0][1] = -1.0;
    d[1][0] = 1.0;  d[1][1] = 0.0;
    d[2][0] = 0.0;  d[2][1] = 1.0;
}

/* P2 reference basis on the triangle (src/fedefs/h1_p2.jl:134-145): nodes then
 * faces (1-2), (2-3), (3-1). */
static void p2_tri_vals(const double *x, double *phi)
{
    double l = 1.0 - x[0] - x[1];
    phi[0] = 2.0 * l * (l - 0.5);
    phi[1] = 2.0 * x[0] * (x[0] - 0.5);
    phi[2] = 2.0 * x[1] * (x[1] - 0.5);
    phi[3] = 4.0 * l * x[0];
    phi[4] = 4.0 * x[0] * x[1];
    phi[5] = 4.0 * x[1] * l;
}
/* Exact derivatives of the P2 basis.  At the dyadic points of the order-2
 * rule every value below is exact, so it equals what forward-mode AD of the
 * basis expressions produces there (src/feevaluator.jl:248-290). */
static void p2_tri_ders(const double *x, double d[6][2])
{
    double l = 1.0 - x[0] - x[1];
    d[0][0] = -(4.0 * l - 1.0); d[0][1] = -(4.0 * l - 1.0);
    d[1][0] = 4.0 * x[0] - 1.0; d[1][1] = 0.0;
    d[2][0] = 0.0;              d[2][1] = 4.0 * x[1] - 1.0;
    d[3][0] = 4.0 * (l - x[0]); d[3][1] = -4.0 * x[0];
    d[4][0] = 4.0 * x[1];       d[4][1] = 4.0 * x[0];
    d[5][0] = -4.0 * x[1];      d[5][1] = 4.0 * (l - x[1]);
}

/* ---------------------------------------------------------- affine maps */

/* Triangle: A = [x2-x1, x3-x1] (columns), det = A11 A22 - A12 A21,
 * M = A^{-T} = [A22 -A21; -A12 A11]/det, vol = |det|/2. */
static void trafo_tri(const double *X, double M[2][2], double *vol)
{
    double A11 = X[2] - X[0], A12 = X[4] - X[0];
    double A21 = X[3] - X[1], A22 = X[5] - X[1];
    double det = A11 * A22 - A12 * A21;
    M[0][0] = A22 / det;
    M[0][1] = -A21 / det;
    M[1][0] = -A12 / det;
    M[1][1] = A11 / det;
    *vol = fabs(det) * 0.5;
}

/* Tetrahedron: A columns x_{k+1}-x_1; M = A^{-T} from cofactors; vol=|det|/6. */
static void trafo_tet(const double *X, double M[3][3], double *vol)
{
    double A[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) A[r][k] = X[3 * (k + 1) + r] - X[r];
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
    /* (A^{-1})_{kr} = cof_{rk}/det, M = A^{-T} so M_{rk} = cof_{rk}/det */
    M[0][0] = c00 / det; M[0][1] = c01 / det; M[0][2] = c02 / det;
    M[1][0] = c10 / det; M[1][1] = c11 / det; M[1][2] = c12 / det;
    M[2][0] = c20 / det; M[2][1] = c21 / det; M[2][2] = c22 / det;
    *vol = fabs(det) / 6.0;
}

/* ------------------------------------------- extendable sparse (restated) */

typedef struct {
    int64_t n;          /* columns */
    int32_t *cnt, *cap; /* per column */
    int32_t **row;
    double **val;
} or_lnk;

static or_lnk *lnk_new(int64_t n)
{
    or_lnk *L = (or_lnk *)calloc(1, sizeof(or_lnk));
    L->n = n;
    L->cnt = (int32_t *)calloc(n, sizeof(int32_t));
    L->cap = (int32_t *)calloc(n, sizeof(int32_t));
    L->row = (int32_t **)calloc(n, sizeof(int32_t *));
    L->val = (double **)calloc(n, sizeof(double *));
    return L;
}

/* rawupdateindex!(A, +, v, i, j): add to an existing entry, or create it with
 * value 0 + v.  Caller applies the _addnz zero test first. */
static void lnk_add(or_lnk *L, int32_t i, int32_t j, double v)
{
    int32_t n = L->cnt[j];
    int32_t *r = L->row[j];
    for (int32_t p = 0; p < n; ++p)
        if (r[p] == i) { L->val[j][p] += v; return; }
    if (n == L->cap[j]) {
        int32_t nc = L->cap[j] ? 2 * L->cap[j] : 8;
        L->row[j] = (int32_t *)realloc(L->row[j], sizeof(int32_t) * nc);
        L->val[j] = (double *)realloc(L->val[j], sizeof(double) * nc);
        L->cap[j] = nc;
    }
    L->row[j][n] = i;
    L->val[j][n] = 0.0 + v;
    L->cnt[j] = n + 1;
}

/* setindex! used by apply_penalties!: overwrite, inserting if absent. */
static void lnk_set(or_lnk *L, int32_t i, int32_t j, double v)
{
    int32_t n = L->cnt[j];
    for (int32_t p = 0; p < n; ++p)
        if (L->row[j][p] == i) { L->val[j][p] = v; return; }
    lnk_add(L, i, j, 0.0);
    L->val[j][L->cnt[j] - 1] = v;
}

/* ----------------------------------------------------------- matrix handle */

typedef struct {
    int64_t n, nnz;
    or_lnk *L;
    int64_t *colptr; /* 0-based CSC after flush */
    int32_t *rowval;
    double *nzval;
} or_matrix;

void *or_matrix_new(int64_t n)
{
    or_matrix *A = (or_matrix *)calloc(1, sizeof(or_matrix));
    A->n = n;
    A->L = lnk_new(n);
    return A;
}

void or_matrix_free(void *h)
{
    or_matrix *A = (or_matrix *)h;
    if (!A) return;
    for (int64_t j = 0; j < A->n; ++j) { free(A->L->row[j]); free(A->L->val[j]); }
    free(A->L->cnt); free(A->L->cap); free(A->L->row); free(A->L->val); free(A->L);
    free(A->colptr); free(A->rowval); free(A->nzval);
    free(A);
}

static int cmp_rv(const void *a, const void *b)
{
    const int32_t ra = *(const int32_t *)a, rb = *(const int32_t *)b;
    return (ra > rb) - (ra < rb);
}

/* flush!: CSC with rows sorted per column.  Returns nnz. */
int64_t or_matrix_flush(void *h)
{
    or_matrix *A = (or_matrix *)h;
    free(A->colptr); free(A->rowval); free(A->nzval);
    int64_t nnz = 0;
    for (int64_t j = 0; j < A->n; ++j) nnz += A->L->cnt[j];
    A->nnz = nnz;
    A->colptr = (int64_t *)malloc(sizeof(int64_t) * (A->n + 1));
    A->rowval = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    A->nzval = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    int64_t p = 0;
    struct rv { int32_t r; int32_t pad; double v; };
    struct rv *tmp = NULL; int32_t tcap = 0;
    for (int64_t j = 0; j < A->n; ++j) {
        A->colptr[j] = p;
        int32_t c = A->L->cnt[j];
        if (c > tcap) { tcap = c; tmp = (struct rv *)realloc(tmp, sizeof(struct rv) * tcap); }
        for (int32_t q = 0; q < c; ++q) { tmp[q].r = A->L->row[j][q]; tmp[q].v = A->L->val[j][q]; }
        qsort(tmp, c, sizeof(struct rv), cmp_rv);
        for (int32_t q = 0; q < c; ++q) { A->rowval[p] = tmp[q].r; A->nzval[p] = tmp[q].v; ++p; }
    }
    A->colptr[A->n] = p;
    free(tmp);
    return nnz;
}

void or_matrix_export(void *h, int64_t *colptr, int32_t *rowval, double *nzval)
{
    or_matrix *A = (or_matrix *)h;
    memcpy(colptr, A->colptr, sizeof(int64_t) * (A->n + 1));
    memcpy(rowval, A->rowval, sizeof(int32_t) * A->nnz);
    memcpy(nzval, A->nzval, sizeof(double) * A->nnz);
}

/* apply_penalties!: A[d,d] = penalty for each fixed dof (src/fematrix.jl:349-355). */
void or_matrix_penalize(void *h, int64_t nfixed, const int32_t *fixed, double penalty)
{
    or_matrix *A = (or_matrix *)h;
    for (int64_t k = 0; k < nfixed; ++k) lnk_set(A->L, fixed[k], fixed[k], penalty);
}

/* -------------------------------------------------- element operator values */

/* Evaluates cvals[resultdim][ndofs][nq] for one cell, in the accumulation
 * order of src/feevaluator_h1.jl (Gradient :61-74, SymmetricGradient :97-116).
 * geometry: 2 = triangle, 3 = tetrahedron.  Returns ndofs. */
typedef struct {
    int geom, fe, ncomp, op;   /* op: 1 Identity, 2 Gradient, 3 SymmetricGradient */
    int nq;
    double xref[16][3], w[16];
    int nsd;                   /* scalar dofs per component */
    double D[16][10][3];       /* [qp][scalar dof][ref dir] */
    double V[16][10];          /* [qp][scalar dof] */
    int resultdim, ndofs;
} or_eval;

static void eval_setup(or_eval *E, int geom, int fe, int ncomp, int op, int quadorder)
{
    memset(E, 0, sizeof(*E));
    E->geom = geom; E->fe = fe; E->ncomp = ncomp; E->op = op;
    if (geom == 2) {
        double xr[16 * 2];
        E->nq = or_quad_tri(quadorder, xr, E->w);
        for (int q = 0; q < E->nq; ++q) { E->xref[q][0] = xr[2 * q]; E->xref[q][1] = xr[2 * q + 1]; }
        for (int q = 0; q < E->nq; ++q) {
            if (fe == OR_FE_P1) {
                double d[3][2]; p1_tri_ders(d); p1_tri_vals(E->xref[q], E->V[q]);
                for (int a = 0; a < 3; ++a) { E->D[q][a][0] = d[a][0]; E->D[q][a][1] = d[a][1]; }
                E->nsd = 3;
            } else if (fe == OR_FE_P2) {
                double d[6][2]; p2_tri_ders(E->xref[q], d); p2_tri_vals(E->xref[q], E->V[q]);
                for (int a = 0; a < 6; ++a) { E->D[q][a][0] = d[a][0]; E->D[q][a][1] = d[a][1]; }
                E->nsd = 6;
            }
        }
    } else {
        /* Tetrahedron3D order <= 1: midpoint (src/quadrature.jl:276-280) */
        E->nq = 1;
        double t = 1.0 / 4.0;
        E->xref[0][0] = E->xref[0][1] = E->xref[0][2] = t; E->w[0] = 1.0;
        double l = 1.0; l -= t; l -= t; l -= t;
        E->V[0][0] = l; E->V[0][1] = t; E->V[0][2] = t; E->V[0][3] = t;
        static const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int a = 0; a < 4; ++a) for (int j = 0; j < 3; ++j) E->D[0][a][j] = d[a][j];
        E->nsd = 4;
    }
    int edim = geom;
    E->ndofs = E->nsd * ncomp;
    if (op == 1) E->resultdim = ncomp;
    else if (op == 2) E->resultdim = edim * ncomp;
    else E->resultdim = (edim == 2) ? 3 : 6;
}

/* cv[q][dof][k] */
static void eval_cell(const or_eval *E, const double *M /*edim x edim row-major*/, double cv[16][20][6])
{
    int edim = E->geom;
    for (int q = 0; q < E->nq; ++q)
        for (int dof = 0; dof < E->ndofs; ++dof)
            for (int k = 0; k < E->resultdim; ++k) cv[q][dof][k] = 0.0;
    if (E->op == 1) {
        for (int q = 0; q < E->nq; ++q)
            for (int c = 0; c < E->ncomp; ++c)
                for (int a = 0; a < E->nsd; ++a) cv[q][c * E->nsd + a][c] = E->V[q][a];
        return;
    }
    if (E->op == 2) {
        /* for i, c, j, k, dof: cvals[k+off[c],dof,i] += Ainv[k,j]*D[dof+off2[c],j,i] */
        for (int q = 0; q < E->nq; ++q)
            for (int c = 0; c < E->ncomp; ++c)
                for (int j = 0; j < edim; ++j)
                    for (int k = 0; k < edim; ++k)
                        for (int dof = 0; dof < E->ndofs; ++dof) {
                            int cc = dof / E->nsd, a = dof % E->nsd;
                            double d = (cc == c) ? E->D[q][a][j] : 0.0;
                            cv[q][dof][k + c * edim] += M[k * edim + j] * d;
                        }
        return;
    }
    /* SymmetricGradient{1}: for i, dof, c, k, j; compression [1,3,3,2] (2D) */
    static const int comp2[4] = {0, 2, 2, 1};
    static const int comp3[9] = {0, 5, 4, 5, 1, 3, 4, 3, 2};
    for (int q = 0; q < E->nq; ++q)
        for (int dof = 0; dof < E->ndofs; ++dof)
            for (int c = 0; c < E->ncomp; ++c)
                for (int k = 0; k < edim; ++k)
                    for (int j = 0; j < edim; ++j) {
                        int cc = dof / E->nsd, a = dof % E->nsd;
                        double d = (cc == c) ? E->D[q][a][j] : 0.0;
                        int t = (edim == 2) ? comp2[k + c * edim] : comp3[k + c * edim];
                        if (k != c) cv[q][dof][t] += 1.0 * M[k * edim + j] * d;
                        else cv[q][dof][t] += M[k * edim + j] * d;
                    }
}

static void cell_coords(int geom, const double *coords, const int32_t *cells, int64_t c, double *X)
{
    int nv = geom + 1;
    for (int a = 0; a < nv; ++a)
        for (int r = 0; r < geom; ++r) X[geom * a + r] = coords[(int64_t)geom * cells[(int64_t)nv * c + a] + r];
}

/* Hooke tensor action (src/pdeoperators.jl:256-276) with per-cell scaling:
 * the Lame parameters of cell c are (kappa_c*lambda, kappa_c*mu). */
static void hooke2d(double lam, double mu, const double *in, double *out)
{
    out[0] = (lam + 2.0 * mu) * in[0] + lam * in[1];
    out[1] = (lam + 2.0 * mu) * in[1] + lam * in[0];
    out[2] = mu * in[2];
}

/* ----------------------------------------------------- BilinearForm loop */

/* Assembles one BilinearForm over all cells into matrix h (row offset ro,
 * column offset co) following src/assemblypatterns/bilinearform.jl:226-377.
 *   kind: OR_BLF_LAPLACE / OR_BLF_MASS (symmetric pattern type) or
 *         OR_BLF_HOOKE2D (full pattern type, action on the first argument)
 *   coef_mode: OR_COEF_FACTOR -> constant coefficient passed as factor
 *              (NoAction; enters itemfactor = vol*factor*1);
 *              OR_COEF_CELL  -> kappa[c] enters through the action
 *              (result = kappa[c] .* input); factor is the operator factor
 *              (1 for BilinearForm(...; action)).
 *   params: Hooke (lambda, mu). */
void or_assemble_blf(void *h, int kind, int geom, int fe, int ncomp, int quadorder,
                     int coef_mode, double factor, const double *kappa, const double *params,
                     int64_t nc, const double *coords, const int32_t *cells,
                     const int32_t *celldofs, int32_t ro, int32_t co)
{
    or_matrix *A = (or_matrix *)h;
    or_eval E;
    int op = (kind == OR_BLF_HOOKE2D) ? 3 : (kind == OR_BLF_MASS ? 1 : 2);
    eval_setup(&E, geom, fe, ncomp, op, quadorder);
    int nd = E.ndofs, rd = E.resultdim;
    int symmetric = (kind != OR_BLF_HOOKE2D);
    double X[12], M[9], vol, L[20][20], cv[16][20][6], ar[6];
    for (int64_t c = 0; c < nc; ++c) {
        cell_coords(geom, coords, cells, c, X);
        if (geom == 2) { double m[2][2]; trafo_tri(X, m, &vol); M[0] = m[0][0]; M[1] = m[0][1]; M[2] = m[1][0]; M[3] = m[1][1]; }
        else { double m[3][3]; trafo_tet(X, m, &vol); for (int r = 0; r < 9; ++r) M[r] = m[r / 3][r % 3]; }
        eval_cell(&E, M, cv);
        for (int a = 0; a < nd; ++a) for (int b = 0; b < nd; ++b) L[a][b] = 0.0;
        for (int q = 0; q < E.nq; ++q) {
            for (int a = 0; a < nd; ++a) {
                if (kind == OR_BLF_HOOKE2D) {
                    double kc = (coef_mode == OR_COEF_CELL) ? kappa[c] : 1.0;
                    hooke2d(kc * params[0], kc * params[1], cv[q][a], ar);
                } else if (coef_mode == OR_COEF_CELL) {
                    for (int k = 0; k < rd; ++k) ar[k] = kappa[c] * cv[q][a][k];
                } else {
                    for (int k = 0; k < rd; ++k) ar[k] = cv[q][a][k];
                }
                for (int b = symmetric ? a : 0; b < nd; ++b) {
                    double t = 0.0;
                    for (int k = 0; k < rd; ++k) t += ar[k] * cv[q][b][k];
                    L[a][b] += E.w[q] * t;
                }
            }
        }
        double itemfactor = vol * factor * 1.0;
        const int32_t *dofs = celldofs + (int64_t)nd * c;
        if (symmetric) {
            for (int a = 0; a < nd; ++a)
                for (int b = a + 1; b < nd; ++b) {
                    double v = L[a][b] * itemfactor;
                    if (v != 0.0) {
                        lnk_add(A->L, dofs[a] + ro, dofs[b] + co, v * 1.0);
                        lnk_add(A->L, dofs[b] + ro, dofs[a] + co, v * 1.0);
                    }
                }
            for (int a = 0; a < nd; ++a) {
                double v = L[a][a] * itemfactor;
                if (v != 0.0) lnk_add(A->L, dofs[a] + ro, dofs[a] + co, v * 1.0);
            }
        } else {
            for (int a = 0; a < nd; ++a)
                for (int b = 0; b < nd; ++b) {
                    double v = L[a][b] * itemfactor;
                    if (v != 0.0) lnk_add(A->L, dofs[a] + ro, dofs[b] + co, v * 1.0);
                }
        }
    }
}

/* ------------------------------------------------------- LinearForm loop */

/* b += LinearForm(Identity, constant f) over all cells, following
 * src/assemblypatterns/linearform.jl:181-220 with the constant-data action
 * (src/actions.jl:119-128): localb[dof] += (sum_k f_k phi_k) * w;
 * b[dof] += localb * (factor*vol*1). */
void or_assemble_lf_const(double *b, int geom, int fe, int ncomp, int quadorder,
                          const double *f, double factor, int64_t nc, const double *coords,
                          const int32_t *cells, const int32_t *celldofs, int32_t off)
{
    or_eval E;
    eval_setup(&E, geom, fe, ncomp, 1, quadorder);
    int nd = E.ndofs;
    double X[12], M[9], vol, cv[16][20][6], lb[20];
    for (int r = 0; r < 9; ++r) M[r] = 0.0;
    for (int64_t c = 0; c < nc; ++c) {
        cell_coords(geom, coords, cells, c, X);
        if (geom == 2) { double m[2][2]; trafo_tri(X, m, &vol); }
        else { double m[3][3]; trafo_tet(X, m, &vol); }
        eval_cell(&E, M, cv);
        for (int a = 0; a < nd; ++a) lb[a] = 0.0;
        for (int q = 0; q < E.nq; ++q)
            for (int a = 0; a < nd; ++a) {
                double t = 0.0;
                for (int k = 0; k < E.resultdim; ++k) t += f[k] * cv[q][a][k];
                lb[a] += t * E.w[q];
            }
        double itemfactor = factor * vol * 1.0;
        const int32_t *dofs = celldofs + (int64_t)nd * c;
        for (int a = 0; a < nd; ++a) b[dofs[a] + off] += lb[a] * itemfactor;
    }
}

/* ------------------------------------------------------- ItemIntegrator */

/* Integral of an FE function's Identity over each cell (NoAction path of
 * src/assemblypatterns/itemintegrator.jl:276-283): out[c*ncomp+j] +=
 * (sum_dofs coeff*phi) * w * vol.  Used for means and for L2 checks. */
void or_integrate_identity(double *out, int geom, int fe, int ncomp, int quadorder,
                           const double *u, int64_t nc, const double *coords,
                           const int32_t *cells, const int32_t *celldofs)
{
    or_eval E;
    eval_setup(&E, geom, fe, ncomp, 1, quadorder);
    int nd = E.ndofs;
    double X[12], vol, cv[16][20][6], M[9] = {0};
    for (int64_t c = 0; c < nc; ++c) {
        cell_coords(geom, coords, cells, c, X);
        if (geom == 2) { double m[2][2]; trafo_tri(X, m, &vol); }
        else { double m[3][3]; trafo_tet(X, m, &vol); }
        eval_cell(&E, M, cv);
        const int32_t *dofs = celldofs + (int64_t)nd * c;
        for (int q = 0; q < E.nq; ++q) {
            double in[6] = {0, 0, 0, 0, 0, 0};
            for (int a = 0; a < nd; ++a)
                for (int k = 0; k < E.resultdim; ++k) in[k] += u[dofs[a]] * cv[q][a][k];
            for (int k = 0; k < E.resultdim; ++k) out[c * E.resultdim + k] += in[k] * E.w[q] * vol;
        }
    }
}

/* --------------------------------------------------------------- solver */

/* Jacobi-preconditioned CG in FP64 on the CSC matrix of the penalized system,
 * with fixed dofs eliminated (x_fixed = b_fixed / A_dd, rows and columns of
 * fixed dofs masked).  Symmetric A: CSC of A == CSR of A^T == CSR of A.
 * Stops at ||r|| <= rtol ||b_free||.  Returns iterations (negative if maxit
 * reached).  Used as the CPU baseline solve; threads = OpenMP threads. */
int64_t or_pcg(int64_t n, const int64_t *colptr, const int32_t *rowval, const double *nzval,
               const double *b, double *x, const uint8_t *fixed, double rtol, int64_t maxit,
               int threads, double *relres)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
    double *dinv = (double *)malloc(sizeof(double) * n);
    double bb = 0.0;
#pragma omp parallel for reduction(+ : bb)
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t k = colptr[i]; k < colptr[i + 1]; ++k)
            if (rowval[k] == i) d = nzval[k];
        if (fixed[i]) { dinv[i] = 0.0; x[i] = b[i] / d; }
        else { dinv[i] = 1.0 / d; x[i] = 0.0; bb += b[i] * b[i]; }
    }
    /* r = b - A x over free rows, using only fixed columns (x_free = 0) */
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        if (fixed[i]) { r[i] = 0.0; continue; }
        double s = 0.0;
        for (int64_t k = colptr[i]; k < colptr[i + 1]; ++k) s += nzval[k] * x[rowval[k]];
        r[i] = b[i] - s;
    }
    double rz = 0.0, rr = 0.0;
#pragma omp parallel for reduction(+ : rz, rr)
    for (int64_t i = 0; i < n; ++i) { z[i] = dinv[i] * r[i]; p[i] = z[i]; rz += r[i] * z[i]; rr += r[i] * r[i]; }
    double bn = sqrt(bb > 0 ? bb : 1.0);
    int64_t it = 0;
    while (sqrt(rr) > rtol * bn && it < maxit) {
        double pq = 0.0;
#pragma omp parallel for reduction(+ : pq)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            if (!fixed[i])
                for (int64_t k = colptr[i]; k < colptr[i + 1]; ++k) {
                    int32_t j = rowval[k];
                    if (!fixed[j]) s += nzval[k] * p[j];
                }
            q[i] = s; pq += p[i] * s;
        }
        double alpha = rz / pq, rz2 = 0.0;
        rr = 0.0;
#pragma omp parallel for reduction(+ : rz2, rr)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i]; r[i] -= alpha * q[i]; z[i] = dinv[i] * r[i];
            rz2 += r[i] * z[i]; rr += r[i] * r[i];
        }
        double beta = rz2 / rz; rz = rz2;
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    *relres = sqrt(rr) / bn;
    free(r); free(z); free(p); free(q); free(dinv);
    return (sqrt(rr) > rtol * bn) ? -it : it;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ============================================================== Stokes ===
 * Bernardi-Raugel H1BR{2} x L2P0 with the BR -> RT0 reconstructed right-hand
 * side, following:
 *   H1BR{2} basis / dofmap "N1f1" / face-normal coefficients ... src/fedefs/h1v_br.jl:15-35,117-130,150-162
 *   Gradient with coefficients .................................. src/feevaluator_h1.jl:77-94
 *   Divergence with coefficients ................................ src/feevaluator_h1.jl:129-145
 *   LagrangeMultiplier(Divergence) + transposed copy ............ src/pdeoperators.jl:214-226,
 *                                                                  src/assemblypatterns/bilinearform.jl:348-364
 *   HDIVRT0{2} basis, CellFaceSigns coefficients ................ src/fedefs/hdiv_rt0.jl:67-73,106-116
 *   Piola map of ReconstructionIdentity ......................... src/feevaluator_h1.jl:342-381
 *   BR -> RT0 reconstruction coefficients ....................... src/reconstructions.jl:381-403
 *   Stroud conical product rule (Golub-Welsch) .................. src/quadrature.jl:528-562
 *   FixedIntegralMean ........................................... src/globalconstraints.jl:259-283,337-351
 * FaceNormals/FaceVolumes/CellFaceSigns come from ExtendableGrids (not in
 * the reference tree): restated as n = (t_y, -t_x)/|t| for the face tangent
 * t = x_b - x_a in its global orientation (a, b) (first cell that met it), and
 * sign = +1 where the cell's local face orientation equals the global one. */

void or_face_geometry(int64_t nc, int64_t nf, const double *coords, const int32_t *cells, const int32_t *cellfaces,
                      const int32_t *facenodes, double *fnormal /*[nf][2]*/, double *fvol, int8_t *csign /*[nc][3]*/)
{
    for (int64_t f = 0; f < nf; ++f) {
        int32_t a = facenodes[2 * f], b = facenodes[2 * f + 1];
        double tx = coords[2 * b] - coords[2 * a], ty = coords[2 * b + 1] - coords[2 * a + 1];
        double len = sqrt(tx * tx + ty * ty);
        fvol[f] = len;
        fnormal[2 * f] = ty / len;
        fnormal[2 * f + 1] = -tx / len;
    }
    for (int64_t c = 0; c < nc; ++c)
        for (int k = 0; k < 3; ++k) {
            int32_t f = cellfaces[3 * c + k];
            int32_t a = cells[3 * c + k], b = cells[3 * c + (k + 1) % 3];
            csign[3 * c + k] = (facenodes[2 * f] == a && facenodes[2 * f + 1] == b) ? 1 : -1;
        }
}

/* "N1f1": component-blocked node dofs (offset nnodes), then one dof per face */
void or_br_celldofs(int64_t nc, int64_t nnodes, const int32_t *cells, const int32_t *cellfaces, int32_t *celldofs)
{
    for (int64_t c = 0; c < nc; ++c) {
        for (int a = 0; a < 3; ++a) {
            celldofs[9 * c + a] = cells[3 * c + a];
            celldofs[9 * c + 3 + a] = (int32_t)(cells[3 * c + a] + nnodes);
            celldofs[9 * c + 6 + a] = (int32_t)(2 * nnodes + cellfaces[3 * c + a]);
        }
    }
}

/* reference gradients of the BR scalar shape functions: P1 nodes 0..2, face
 * bubbles 3..5 (6 x1 l1, 6 x1 x2, 6 l1 x2 with l1 = 1 - x1 - x2) */
static void br_ref_ders(const double *x, double d[6][2])
{
    double l = 1.0;
    l -= x[0];
    l -= x[1];
    p1_tri_ders(d);
    d[3][0] = 6.0 * (l - x[0]); d[3][1] = -6.0 * x[0];
    d[4][0] = 6.0 * x[1];       d[4][1] = 6.0 * x[0];
    d[5][0] = -6.0 * x[1];      d[5][1] = 6.0 * (l - x[1]);
}

/* cvals of Gradient (resultdim 4) and Divergence (resultdim 1) of the 9 BR
 * dofs at one point, in the reference's accumulation order */
static void br_cvals(const double M[2][2], const double *x, const double n[3][2], double grad[9][4], double div[9])
{
    double d[6][2];
    br_ref_ders(x, d);
    for (int dof = 0; dof < 9; ++dof) {
        /* scalar function and coefficient per component */
        int sf = dof < 6 ? dof % 3 : 3 + (dof - 6);
        for (int c = 0; c < 2; ++c) {
            int active = dof < 6 ? (dof / 3 == c) : 1;
            double coef = dof < 6 ? 1.0 : n[dof - 6][c];
            for (int k = 0; k < 2; ++k) {
                double s = 0.0;
                for (int j = 0; j < 2; ++j) s += M[k][j] * (active ? d[sf][j] : 0.0);
                s *= coef;
                grad[dof][k + 2 * c] = s;
            }
        }
        double dv = 0.0;
        for (int k = 0; k < 2; ++k) {
            int active = dof < 6 ? (dof / 3 == k) : 1;
            double coef = dof < 6 ? 1.0 : n[dof - 6][k];
            for (int j = 0; j < 2; ++j) dv += M[k][j] * (active ? d[sf][j] : 0.0) * coef;
        }
        div[dof] = dv;
    }
}

static void cell_face_normals(int64_t c, const int32_t *cellfaces, const double *fnormal, double n[3][2])
{
    for (int k = 0; k < 3; ++k) {
        int32_t f = cellfaces[3 * c + k];
        n[k][0] = fnormal[2 * f];
        n[k][1] = fnormal[2 * f + 1];
    }
}

/* Stokes system blocks into matrix h: [1,1] = LaplaceOperator(nu) on BR
 * (symmetric, quadorder 2), [1,2] = LagrangeMultiplier(Divergence) with
 * factor -1 (quadorder 1) and its transposed copy into [2,1]. */
void or_assemble_stokes(void *h, double nu, int64_t nc, int64_t nnodes, int64_t nf, const double *coords,
                        const int32_t *cells, const int32_t *cellfaces, const double *fnormal)
{
    or_matrix *A = (or_matrix *)h;
    const int64_t nU = 2 * nnodes + nf;
    double X[6], M[2][2], vol;
    /* [1,1] */
    double xr[6], w[3];
    int nq = or_quad_tri(2, xr, w);
    for (int64_t c = 0; c < nc; ++c) {
        cell_coords(2, coords, cells, c, X);
        trafo_tri(X, M, &vol);
        double n[3][2];
        cell_face_normals(c, cellfaces, fnormal, n);
        double G[3][9][4], D[3][9], L[9][9];
        for (int q = 0; q < nq; ++q) br_cvals(M, xr + 2 * q, n, G[q], D[q]);
        for (int a = 0; a < 9; ++a)
            for (int b = 0; b < 9; ++b) L[a][b] = 0.0;
        for (int q = 0; q < nq; ++q)
            for (int a = 0; a < 9; ++a)
                for (int b = a; b < 9; ++b) {
                    double t = 0.0;
                    for (int k = 0; k < 4; ++k) t += G[q][a][k] * G[q][b][k];
                    L[a][b] += w[q] * t;
                }
        double itemfactor = vol * nu * 1.0;
        int32_t dofs[9];
        for (int a = 0; a < 3; ++a) {
            dofs[a] = cells[3 * c + a];
            dofs[3 + a] = (int32_t)(cells[3 * c + a] + nnodes);
            dofs[6 + a] = (int32_t)(2 * nnodes + cellfaces[3 * c + a]);
        }
        for (int a = 0; a < 9; ++a)
            for (int b = a + 1; b < 9; ++b) {
                double v = L[a][b] * itemfactor;
                if (v != 0.0) {
                    lnk_add(A->L, dofs[a], dofs[b], v * 1.0);
                    lnk_add(A->L, dofs[b], dofs[a], v * 1.0);
                }
            }
        for (int a = 0; a < 9; ++a) {
            double v = L[a][a] * itemfactor;
            if (v != 0.0) lnk_add(A->L, dofs[a], dofs[a], v * 1.0);
        }
    }
    /* [1,2] and its transposed copy */
    nq = or_quad_tri(1, xr, w);
    const double factor = -1.0, factor_t = -1.0;
    for (int64_t c = 0; c < nc; ++c) {
        cell_coords(2, coords, cells, c, X);
        trafo_tri(X, M, &vol);
        double n[3][2];
        cell_face_normals(c, cellfaces, fnormal, n);
        double G[9][4], D[9], L[9];
        br_cvals(M, xr, n, G, D);
        for (int a = 0; a < 9; ++a) {
            double t = 0.0;
            t += D[a] * 1.0;
            L[a] = 0.0;
            L[a] += w[0] * t;
        }
        double itemfactor = vol * factor * 1.0;
        for (int a = 0; a < 3; ++a) (void)a;
        int32_t dofs[9];
        for (int a = 0; a < 3; ++a) {
            dofs[a] = cells[3 * c + a];
            dofs[3 + a] = (int32_t)(cells[3 * c + a] + nnodes);
            dofs[6 + a] = (int32_t)(2 * nnodes + cellfaces[3 * c + a]);
        }
        int32_t prow = (int32_t)(nU + c);
        for (int a = 0; a < 9; ++a) {
            double v = L[a] * itemfactor;
            if (v != 0.0) lnk_add(A->L, dofs[a], prow, v * 1.0);
            double vt = L[a] * itemfactor / factor * factor_t;
            if (vt != 0.0) lnk_add(A->L, prow, dofs[a], vt * -1.0);
        }
    }
}

/* ---- Stroud rule: Gauss-Legendre x Gauss-Jacobi(1,0) by Golub-Welsch,
 * eigenpairs of the symmetric tridiagonal Jacobi matrices by cyclic Jacobi
 * rotations, then the conical product (src/quadrature.jl:528-562). */
static void sym_eigen(int n, double *Amat /*n x n, destroyed*/, double *evals, double *evecs /*n x n, column k = vector k*/)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) evecs[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += Amat[i * n + j] * Amat[i * n + j];
        if (off < 1e-60) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = Amat[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double theta = (Amat[q * n + q] - Amat[p * n + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    double akp = Amat[k * n + p], akq = Amat[k * n + q];
                    Amat[k * n + p] = cs * akp - sn * akq;
                    Amat[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = Amat[p * n + k], aqk = Amat[q * n + k];
                    Amat[p * n + k] = cs * apk - sn * aqk;
                    Amat[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = evecs[k * n + p], vkq = evecs[k * n + q];
                    evecs[k * n + p] = cs * vkp - sn * vkq;
                    evecs[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) evals[i] = Amat[i * n + i];
    /* ascending order */
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (evals[j] < evals[i]) {
                double t = evals[i]; evals[i] = evals[j]; evals[j] = t;
                for (int k = 0; k < n; ++k) {
                    double v = evecs[k * n + i]; evecs[k * n + i] = evecs[k * n + j]; evecs[k * n + j] = v;
                }
            }
}

int or_quad_stroud(int order, double *xref /*[nq][2]*/, double *w)
{
    int n = order / 2 + 1;
    double J[64], ev[8], V[64], r[8], a[8], s[8], b[8];
    for (int i = 0; i < n * n; ++i) J[i] = 0.0;
    for (int k = 1; k < n; ++k) {
        double g = (double)k / sqrt(4.0 * k * k - 1.0);
        J[(k - 1) * n + k] = J[k * n + k - 1] = g;
    }
    sym_eigen(n, J, ev, V);
    for (int i = 0; i < n; ++i) { r[i] = 0.5 * ev[i] + 0.5; a[i] = 0.5 * (2.0 * V[i] * V[i]); }
    for (int i = 0; i < n * n; ++i) J[i] = 0.0;
    for (int k = 1; k <= n; ++k) J[(k - 1) * n + k - 1] = -1.0 / (4.0 * k * k - 1.0);
    for (int k = 1; k < n; ++k) {
        double g = sqrt((double)(k + 1) * k) / (2.0 * (k + 1) - 1.0);
        J[(k - 1) * n + k] = J[k * n + k - 1] = g;
    }
    sym_eigen(n, J, ev, V);
    for (int i = 0; i < n; ++i) { s[i] = 0.5 * ev[i] + 0.5; b[i] = 0.5 * (2.0 * V[i] * V[i]); }
    /* xref[j*n + i] = (s_j, r_i (1 - s_j)), w = a_i b_j */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            xref[2 * (j * n + i)] = s[j];
            xref[2 * (j * n + i) + 1] = r[i] * (1.0 - s[j]);
            w[j * n + i] = a[i] * b[j];
        }
    return n * n;
}

/* body force f = grad(phi) + curl(psi) with phi = sum_m A_m sin(2 pi (p_m x + q_m y) + t_m)
 * + c0 x^3 + c1 x y^2 + c2 y^2, psi = sum_m B_m sin(2 pi (p'_m x + q'_m y) + t'_m);
 * modes: [K][4] = (amplitude, p, q, phase) for phi, then for psi. */
static void stokes_force(int K, const double *modes, const double *poly, const double *x, double *f)
{
    const double tp = 6.283185307179586;
    double fx = 0.0, fy = 0.0;
    for (int m = 0; m < K; ++m) {
        const double *md = modes + 4 * m;
        double arg = tp * (md[1] * x[0] + md[2] * x[1]) + md[3];
        double cs = md[0] * tp * cos(arg);
        fx += cs * md[1];
        fy += cs * md[2];
    }
    for (int m = 0; m < K; ++m) {
        const double *md = modes + 4 * (K + m);
        double arg = tp * (md[1] * x[0] + md[2] * x[1]) + md[3];
        double cs = md[0] * tp * cos(arg);
        fx += cs * md[2];   /* d psi / dy */
        fy -= cs * md[1];   /* -d psi / dx */
    }
    fx += 3.0 * poly[0] * x[0] * x[0] + poly[1] * x[1] * x[1];
    fy += 2.0 * poly[1] * x[0] * x[1] + 2.0 * poly[2] * x[1];
    f[0] = fx;
    f[1] = fy;
}

/* b[0:nU] += LinearForm(ReconstructionIdentity{HDIVRT0{2}}, f), quadrature
 * order `order` (Stroud for order > 2) */
void or_assemble_stokes_rhs(double *b, int order, int K, const double *modes, const double *poly, int64_t nc,
                            int64_t nnodes, const double *coords, const int32_t *cells, const int32_t *cellfaces,
                            const double *fnormal, const double *fvol, const int8_t *csign)
{
    double xr[2 * 64], w[64];
    int nq = order <= 2 ? or_quad_tri(order, xr, w) : or_quad_stroud(order, xr, w);
    double X[6];
    for (int64_t c = 0; c < nc; ++c) {
        cell_coords(2, coords, cells, c, X);
        double A11 = X[2] - X[0], A12 = X[4] - X[0], A21 = X[3] - X[1], A22 = X[5] - X[1];
        double det = A11 * A22 - A12 * A21;
        double vol = fabs(det) * 0.5;
        double rc[9][3];
        for (int i = 0; i < 9; ++i) for (int j = 0; j < 3; ++j) rc[i][j] = 0.0;
        for (int f = 0; f < 3; ++f) {
            int32_t face = cellfaces[3 * c + f];
            for (int nn = 0; nn < 2; ++nn) {
                int node = (f + nn) % 3;
                for (int k = 0; k < 2; ++k) rc[3 * k + node][f] = 0.5 * fvol[face] * fnormal[2 * face + k];
            }
            rc[6 + f][f] = fvol[face];
        }
        double lb[9];
        for (int i = 0; i < 9; ++i) lb[i] = 0.0;
        for (int q = 0; q < nq; ++q) {
            const double *x = xr + 2 * q;
            /* RT0 reference basis and Piola map with CellFaceSigns */
            double phi[3][2] = {{x[0], x[1] - 1.0}, {x[0], x[1]}, {x[0] - 1.0, x[1]}};
            double T[3][2];
            for (int j = 0; j < 3; ++j) {
                double s = (double)csign[3 * c + j];
                double t0 = 0.0, t1 = 0.0;
                t0 += A11 * phi[j][0];
                t0 += A12 * phi[j][1];
                t1 += A21 * phi[j][0];
                t1 += A22 * phi[j][1];
                t0 *= s / det;
                t1 *= s / det;
                T[j][0] = t0;
                T[j][1] = t1;
            }
            double cv[9][2];
            for (int i = 0; i < 9; ++i) { cv[i][0] = 0.0; cv[i][1] = 0.0; }
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 3; ++j)
                    if (rc[i][j] != 0.0)
                        for (int k = 0; k < 2; ++k) cv[i][k] += rc[i][j] * T[j][k];
            /* physical point x = x1 + A xref */
            double xp[2];
            xp[0] = X[0] + A11 * x[0] + A12 * x[1];
            xp[1] = X[1] + A21 * x[0] + A22 * x[1];
            double f[2];
            stokes_force(K, modes, poly, xp, f);
            for (int i = 0; i < 9; ++i) {
                double t = 0.0;
                for (int k = 0; k < 2; ++k) t += f[k] * cv[i][k];
                lb[i] += t * w[q];
            }
        }
        double itemfactor = 1.0 * vol * 1.0;
        for (int a = 0; a < 3; ++a) {
            b[cells[3 * c + a]] += lb[a] * itemfactor;
            b[cells[3 * c + a] + nnodes] += lb[3 + a] * itemfactor;
            b[2 * nnodes + cellfaces[3 * c + a]] += lb[6 + a] * itemfactor;
        }
    }
}
