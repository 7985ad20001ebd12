/*
 * mhfem_oracle.c -- CPU restatement of the TNL-MHFEM host path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the
 * timed CPU baseline.  Nothing in the product library links or loads it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the TNL-MHFEM tree).  Arithmetic that lives in TNL (not vendored in the
 * reference: face numbering, LU_sequential_*, BiCGStab, GMRES, ILU(0)) is
 * restated from the named published methods; those parts are pinned only by
 * mathematical checks in tests/ (closed forms, quadrature, brute-force mixed
 * systems, scipy), never by a TNL artefact.  Face numbering and the Krylov
 * iteration sequence are "parity unpinned" with respect to TNL itself.
 */
#include "mhfem_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct or_problem {
    int dim, fpc, n, S;
    int shape, nvc; /* OR_SIMPLEX or OR_BOX; vertices per cell */
    int nv, nc, nf;
    int lumping, advection, do_mob;
    double *pts;
    int *cv;
    int *cell_faces, *face_cells;
    double *cell_meas, *face_meas;
    /* model (BaseModel.h:266-286) */
    double *Z_iF, *Z_iK, *N, *u, *m, *D, *w, *a, *r, *f;
    double *v, *m_upw, *Z_upw, *b, *R_ijKe, *R_iK;
    /* boundary conditions (BoundaryConditions.h:76-82) + model boundary mobility */
    uint8_t *tags;
    double *bc_values, *bc_dirichlet, *bc_mobility;
    /* linear system */
    int nrows;
    long nnz;
    int *rowptr, *colidx;
    double *val, *rhs;
    char err[256];
};

/* ------------------------------------------------------------------ */
/* indexing in the reference's CUDA permutations (BaseModel.h:50-124)  */
#define C_ (p->nc)
#define F_ (p->nf)
#define NN (p->n)
#define FP (p->fpc)
#define ZF(j, E) p->Z_iF[(long)(E) * NN + (j)]
#define ZK(i, K) p->Z_iK[(long)(i) * C_ + (K)]
#define IJK(i, j, K) (((long)(i) * NN + (j)) * C_ + (K))
#define IJKE(i, j, K, e) ((((long)(i) * NN + (j)) * FP + (e)) * C_ + (K))
#define IK(i, K) ((long)(i) * C_ + (K))
#define IKE(i, K, e) (((long)(i) * FP + (e)) * C_ + (K))
#define MUPW(i, E) p->m_upw[(long)(E) * NN + (i)]
#define ZUPW(i, j, E) p->Z_upw[((long)(E) * NN + (j)) * NN + (i)]
#define BST(i, j, K, s) p->b[(((long)(i) * NN + (j)) * p->S + (s)) * C_ + (K)]
#define BCI(i, E) ((long)(i) * F_ + (E))

/* Direction of local face e of a box cell, which selects its storage slot:
 * edges one slot; quadrangles e = 0, 2 (horizontal) and e = 1, 3 (vertical);
 * hexahedra n_x faces e = 2, 4, n_y faces e = 1, 3, n_z faces e = 0, 5
 * (MassMatrix.h:178-184, 300-306, 430-436). */
static inline int box_dir(int dim, int e)
{
    static const int hex[6] = {2, 1, 0, 1, 0, 2};
    return dim == 1 ? 0 : dim == 2 ? e % 2 : hex[e];
}

/* MassMatrix::b_ijKef -- simplices: packed upper triangle (MassMatrix.h:613-616,
 * 884-887); boxes: diagonal (lumped) or 2b on the diagonal and b between the
 * two faces of one direction (MassMatrix.h:57-71, 118-132, 190-205, 250-276,
 * 384-399, 444-481) */
static inline double b_ef(const or_problem *p, int i, int j, int K, int e, int f)
{
    if (p->shape == OR_BOX) {
        const int de = box_dir(p->dim, e), df = box_dir(p->dim, f);
        if (p->lumping) return e == f ? BST(i, j, K, de) : 0.0;
        if (de != df) return 0.0;
        return e == f ? 2 * BST(i, j, K, de) : BST(i, j, K, de);
    }
    if (e > f) { int t = e; e = f; f = t; }
    return BST(i, j, K, e + (f * (f + 1)) / 2);
}
/* MassMatrix::b_ijKe -- simplices: row sums (MassMatrix.h:630, 901); boxes:
 * b (lumped) or 3b (MassMatrix.h:73-84, 134-145, 207-221, 278-292, 401-417, 483-499) */
static inline double b_e(const or_problem *p, int i, int j, int K, int e)
{
    if (p->shape == OR_BOX) {
        const double b = BST(i, j, K, box_dir(p->dim, e));
        return p->lumping ? b : 3 * b;
    }
    return BST(i, j, K, p->S - p->fpc + e);
}

/* ------------------------------------------------------------------ */
/* Topology.  Restates TNL's mesh initializer as used through
 * mesh_helpers.h:49-59,151-160,235-241: faces are numbered in the order in
 * which they are first met when cells are visited in index order and the
 * faces of each cell in local order; local face e is the face opposite local
 * vertex e (MassMatrix.h:538,661,787,932); the cells of a face are listed in
 * ascending order.  Parity unpinned w.r.t. TNL (the initializer is not in the
 * reference tree); pinned in tests by invariants and closed-form counts. */
static unsigned long hash3(const int *k, int d)
{
    unsigned long h = 1469598103934665603UL;
    for (int t = 0; t < d; t++) {
        h ^= (unsigned long)(unsigned)k[t];
        h *= 1099511628211UL;
    }
    return h ^ (h >> 29);
}

/* Vertices (positions within the cell) of local face e.  Simplices: every
 * vertex but e.  Boxes, in TNL's subentity order of its topologies: an edge's
 * faces are its vertices; a quadrangle's edges (0,1), (1,2), (2,3), (3,0); a
 * hexahedron's faces (0,1,2,3), (0,1,5,4), (1,2,6,5), (2,3,7,6), (3,0,4,7),
 * (4,5,6,7) (vertices in VTK order).  Returns the count. */
static int face_vertices(int dim, int shape, int e, int *lv)
{
    static const int quad[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static const int hex[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                  {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
    if (shape == OR_SIMPLEX && dim > 1) {
        int t = 0;
        for (int v = 0; v <= dim; v++)
            if (v != e) lv[t++] = v;
        return t;
    }
    if (dim == 1) { lv[0] = e; return 1; }
    if (dim == 2) { lv[0] = quad[e][0]; lv[1] = quad[e][1]; return 2; }
    for (int t = 0; t < 4; t++) lv[t] = hex[e][t];
    return 4;
}

int or_topology(int dim, int nv, const double *points, int nc, const int *cv, int *cell_faces,
                int *face_cells)
{
    return or_topology_shape(dim, dim == 1 ? OR_BOX : OR_SIMPLEX, nv, points, nc, cv, cell_faces,
                             face_cells);
}

int or_topology_shape(int dim, int shape, int nv, const double *points, int nc, const int *cv,
                      int *cell_faces, int *face_cells)
{
    (void)points;
    const int box = shape == OR_BOX || dim == 1;
    const int fpc = box ? 2 * dim : dim + 1, nvc = box ? 1 << dim : dim + 1;
    const int d = box ? 1 << (dim - 1) : dim; /* vertices per face */
    const long maxf = (long)nc * fpc;
    long cap = 1;
    while (cap < 2 * maxf) cap <<= 1;
    int *table = malloc(cap * sizeof(int));
    int *keys = malloc(maxf * d * sizeof(int));
    if (!table || !keys) { free(table); free(keys); return -1; }
    memset(table, 0xff, cap * sizeof(int));
    int nf = 0, bad = 0;
    for (int K = 0; K < nc && !bad; K++) {
        for (int e = 0; e < fpc; e++) {
            int key[4], lv[4];
            face_vertices(dim, box ? OR_BOX : OR_SIMPLEX, e, lv);
            for (int t = 0; t < d; t++) key[t] = cv[(long)K * nvc + lv[t]];
            /* sort the seed (insertion sort, d <= 4) */
            for (int a = 1; a < d; a++)
                for (int c = a; c > 0 && key[c - 1] > key[c]; c--) {
                    int s = key[c]; key[c] = key[c - 1]; key[c - 1] = s;
                }
            for (int a = 0; a < d; a++)
                if (key[a] < 0 || key[a] >= nv) bad = 1;
            unsigned long h = hash3(key, d) & (cap - 1);
            int id = -1;
            while (table[h] >= 0) {
                const int *kk = keys + (long)table[h] * d;
                int eq = 1;
                for (int a = 0; a < d; a++) eq &= (kk[a] == key[a]);
                if (eq) { id = table[h]; break; }
                h = (h + 1) & (cap - 1);
            }
            if (id < 0) {
                id = nf++;
                memcpy(keys + (long)id * d, key, d * sizeof(int));
                table[h] = id;
                face_cells[2L * id] = K;
                face_cells[2L * id + 1] = -1;
            } else {
                if (face_cells[2L * id + 1] != -1) bad = 1; /* more than 2 cells */
                face_cells[2L * id + 1] = K;
            }
            cell_faces[(long)K * fpc + e] = id;
        }
    }
    free(table);
    free(keys);
    return bad ? -1 : nf;
}

/* cell measure (TNL's getEntityMeasure): simplices by the determinant;
 * an edge its length; a quadrangle half the cross product of its diagonals;
 * a hexahedron (a voxel) the product of its edge lengths along the axes */
static double box_measure(int dim, const double *v /* 2^dim points of dim coords */)
{
    if (dim == 1) return fabs(v[1] - v[0]);
    if (dim == 2) {
        const double d1x = v[4] - v[0], d1y = v[5] - v[1], d2x = v[6] - v[2], d2y = v[7] - v[3];
        return 0.5 * fabs(d1x * d2y - d1y * d2x);
    }
    return fabs((v[18] - v[0]) * (v[19] - v[1]) * (v[20] - v[2]));
}

static double simplex_measure(int dim, const double *v /* (dim+1)*dim */)
{
    if (dim == 2) {
        double ax = v[2] - v[0], ay = v[3] - v[1], bx = v[4] - v[0], by = v[5] - v[1];
        return 0.5 * fabs(ax * by - ay * bx);
    }
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; k++) {
        a[k] = v[3 + k] - v[k];
        b[k] = v[6 + k] - v[k];
        c[k] = v[9 + k] - v[k];
    }
    double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                 a[2] * (b[0] * c[1] - b[1] * c[0]);
    return fabs(det) / 6.0;
}

static double face_measure(int dim, const double *v /* dim points of dim coords */)
{
    if (dim == 1) return 1.0; /* a point: the flux through it is not scaled */
    if (dim == 2) {
        double dx = v[2] - v[0], dy = v[3] - v[1];
        return sqrt(dx * dx + dy * dy);
    }
    double a[3], b[3];
    for (int k = 0; k < 3; k++) { a[k] = v[3 + k] - v[k]; b[k] = v[6 + k] - v[k]; }
    double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

/* a hexahedron's face (a planar quadrangle in 3-D): half the cross product
 * of its diagonals */
static double quad_face_measure(const double *v /* 4 points of 3 coords */)
{
    double a[3], b[3];
    for (int k = 0; k < 3; k++) { a[k] = v[6 + k] - v[k]; b[k] = v[9 + k] - v[3 + k]; }
    double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

/* ------------------------------------------------------------------ */
/* LU without pivoting (Doolittle, in place) and the in-place solve: the
 * published method restated for TNL's LU_sequential_factorize /
 * LU_sequential_solve_inplace (MassMatrix.h:571-592, Solver.hpp:849-866). */
static void lu_factorize(double *A, int n)
{
    for (int k = 0; k < n; k++) {
        for (int i = k + 1; i < n; i++) {
            A[i * n + k] /= A[k * n + k];
            for (int j = k + 1; j < n; j++) A[i * n + j] -= A[i * n + k] * A[k * n + j];
        }
    }
}
static void lu_solve(const double *A, int n, double *x)
{
    for (int i = 1; i < n; i++)
        for (int j = 0; j < i; j++) x[i] -= A[i * n + j] * x[j];
    for (int i = n - 1; i >= 0; i--) {
        for (int j = i + 1; j < n; j++) x[i] -= A[i * n + j] * x[j];
        x[i] /= A[i * n + i];
    }
}

/* Geometric local mass matrix M_K of RTN0 (MassMatrix.h:543-569 triangle,
 * :666-695 triangle lumped, :793-831 tetrahedron, :938-978 tetrahedron
 * lumped), its inverse by LU column solves, scaled by D, stored upper-packed
 * column by column, followed by the row sums (:594-597, :864-868).
 * Pinned in tests against exact quadrature of the RTN0 basis. */
void or_local_mass_inverse(int dim, int lumping, const double *P, double measure, double D,
                           double *out)
{
    double M[16];
    const int n = dim + 1;
    if (dim == 2) {
        if (!lumping) {
            double P0[2] = {P[0] - P[4], P[1] - P[5]}, P1[2] = {P[2] - P[4], P[3] - P[5]};
            double P00 = P0[0] * P0[0] + P0[1] * P0[1];
            double P01 = P0[0] * P1[0] + P0[1] * P1[1];
            double P11 = P1[0] * P1[0] + P1[1] * P1[1];
            double den = 24 * measure;
            M[0] = (3 * P00 + P11 - 3 * P01) / den;
            M[4] = (P00 + 3 * P11 - 3 * P01) / den;
            M[8] = (P00 + P11 + P01) / den;
            M[1] = (3 * P01 - P00 - P11) / den;
            M[2] = (P11 - P00 - P01) / den;
            M[5] = (P00 - P11 - P01) / den;
        } else {
            double a01[2] = {P[0] - P[2], P[1] - P[3]}, a02[2] = {P[0] - P[4], P[1] - P[5]},
                   a12[2] = {P[2] - P[4], P[3] - P[5]};
#define D2(x, y) ((x)[0] * (y)[0] + (x)[1] * (y)[1])
            double den = 12 * measure;
            M[0] = (D2(a01, a01) + D2(a02, a02)) / den;
            M[4] = (D2(a01, a01) + D2(a12, a12)) / den;
            M[8] = (D2(a02, a02) + D2(a12, a12)) / den;
            M[1] = (D2(a02, a12)) / den;
            M[2] = (-D2(a01, a12)) / den;
            M[5] = (D2(a01, a02)) / den;
#undef D2
        }
        M[3] = M[1]; M[6] = M[2]; M[7] = M[5];
    } else {
#define D3(x, y) ((x)[0] * (y)[0] + (x)[1] * (y)[1] + (x)[2] * (y)[2])
        if (!lumping) {
            double P0[3], P1[3], P2[3];
            for (int k = 0; k < 3; k++) {
                P0[k] = P[k] - P[9 + k];
                P1[k] = P[3 + k] - P[9 + k];
                P2[k] = P[6 + k] - P[9 + k];
            }
            double P00 = D3(P0, P0), P11 = D3(P1, P1), P22 = D3(P2, P2);
            double P01 = D3(P0, P1), P02 = D3(P0, P2), P12 = D3(P1, P2);
            double den = 180 * measure;
            M[0] = (12 * P00 + 2 * P11 + 2 * P22 - 8 * P01 - 8 * P02 + 2 * P12) / den;
            M[5] = (2 * P00 + 12 * P11 + 2 * P22 - 8 * P01 + 2 * P02 - 8 * P12) / den;
            M[10] = (2 * P00 + 2 * P11 + 12 * P22 + 2 * P01 - 8 * P02 - 8 * P12) / den;
            M[15] = 2 * (P00 + P11 + P22 + P01 + P02 + P12) / den;
            M[1] = (-3 * P00 - 3 * P11 + 2 * P22 + 12 * P01 - 3 * P02 - 3 * P12) / den;
            M[2] = (-3 * P00 + 2 * P11 - 3 * P22 - 3 * P01 + 12 * P02 - 3 * P12) / den;
            M[6] = (2 * P00 - 3 * P11 - 3 * P22 - 3 * P01 - 3 * P02 + 12 * P12) / den;
            M[3] = (-3 * P00 + 2 * P11 + 2 * P22 - 3 * P01 - 3 * P02 + 2 * P12) / den;
            M[7] = (2 * P00 - 3 * P11 + 2 * P22 - 3 * P01 + 2 * P02 - 3 * P12) / den;
            M[11] = (2 * P00 + 2 * P11 - 3 * P22 + 2 * P01 - 3 * P02 - 3 * P12) / den;
        } else {
            double a01[3], a02[3], a03[3], a12[3], a13[3], a23[3];
            for (int k = 0; k < 3; k++) {
                a01[k] = P[k] - P[3 + k];
                a02[k] = P[k] - P[6 + k];
                a03[k] = P[k] - P[9 + k];
                a12[k] = P[3 + k] - P[6 + k];
                a13[k] = P[3 + k] - P[9 + k];
                a23[k] = P[6 + k] - P[9 + k];
            }
            double den = 36 * measure;
            M[0] = (D3(a01, a01) + D3(a02, a02) + D3(a03, a03)) / den;
            M[5] = (D3(a01, a01) + D3(a12, a12) + D3(a13, a13)) / den;
            M[10] = (D3(a02, a02) + D3(a12, a12) + D3(a23, a23)) / den;
            M[15] = (D3(a03, a03) + D3(a13, a13) + D3(a23, a23)) / den;
            M[1] = (D3(a02, a12) + D3(a03, a13)) / den;
            M[2] = (-D3(a01, a12) + D3(a03, a23)) / den;
            M[3] = (-D3(a01, a13) - D3(a02, a23)) / den;
            M[6] = (D3(a01, a02) + D3(a13, a23)) / den;
            M[7] = (D3(a01, a03) - D3(a12, a23)) / den;
            M[11] = (D3(a02, a03) + D3(a12, a13)) / den;
        }
#undef D3
        M[4] = M[1]; M[8] = M[2]; M[9] = M[6]; M[12] = M[3]; M[13] = M[7]; M[14] = M[11];
    }
    lu_factorize(M, n);
    int s = 0;
    for (int c = 0; c < n; c++) {
        double vv[4] = {0, 0, 0, 0};
        vv[c] = 1.0;
        lu_solve(M, n, vv);
        for (int r = 0; r <= c; r++) out[s++] = vv[r] * D;
    }
    /* row sums in ascending f (MassMatrix.h:595-597, 865-868) */
    const int base = s;
    for (int e = 0; e < n; e++) {
        double sum = 0;
        for (int f = 0; f < n; f++) {
            int a = e < f ? e : f, bb = e < f ? f : e;
            sum = (f == 0) ? out[a + bb * (bb + 1) / 2] : sum + out[a + bb * (bb + 1) / 2];
        }
        out[base + e] = sum;
    }
}

/* ------------------------------------------------------------------ */
or_problem *or_create(int dim, int nv, const double *points, int nc, const int *cv, int n,
                      int lumping, int advection, int do_mob)
{
    return or_create_shape(dim, dim == 1 ? OR_BOX : OR_SIMPLEX, nv, points, nc, cv, n, lumping,
                           advection, do_mob);
}

/* Cells: simplices (triangles, tetrahedra) or boxes (edges, axis-aligned
 * rectangles and voxels as Quadrangle and Hexahedron cells, MassMatrix.h:28-508).  b_ijK_storage holds
 * MassMatrix::size values per (i,j,K): 9 / 14 for simplices, dim for boxes. */
or_problem *or_create_shape(int dim, int shape, int nv, const double *points, int nc, const int *cv,
                            int n, int lumping, int advection, int do_mob)
{
    if (dim == 1) shape = OR_BOX;
    if (dim < 1 || dim > 3 || (shape == OR_SIMPLEX && dim == 1)) return NULL;
    or_problem *p = calloc(1, sizeof(or_problem));
    p->dim = dim;
    p->shape = shape;
    p->fpc = shape == OR_BOX ? 2 * dim : dim + 1;
    p->nvc = shape == OR_BOX ? 1 << dim : dim + 1;
    p->S = shape == OR_BOX ? dim : dim == 2 ? 9 : 14;
    p->n = n;
    p->nv = nv;
    p->nc = nc;
    p->lumping = lumping;
    p->advection = advection;
    p->do_mob = do_mob;
    p->pts = malloc(sizeof(double) * nv * dim);
    memcpy(p->pts, points, sizeof(double) * nv * dim);
    p->cv = malloc(sizeof(int) * (long)nc * p->nvc);
    memcpy(p->cv, cv, sizeof(int) * (long)nc * p->nvc);
    p->cell_faces = malloc(sizeof(int) * (long)nc * p->fpc);
    int *fc = malloc(sizeof(int) * 2L * nc * p->fpc);
    int nf = or_topology_shape(dim, shape, nv, points, nc, cv, p->cell_faces, fc);
    if (nf < 0) {
        free(fc);
        or_destroy(p);
        return NULL;
    }
    p->nf = nf;
    p->face_cells = realloc(fc, sizeof(int) * 2L * nf);
    p->cell_meas = malloc(sizeof(double) * nc);
    p->face_meas = malloc(sizeof(double) * nf);
    const int fpc = p->fpc, nvc = p->nvc;
#pragma omp parallel for
    for (int K = 0; K < nc; K++) {
        double vv[24];
        for (int k = 0; k < nvc; k++)
            for (int c = 0; c < dim; c++) vv[k * dim + c] = points[(long)cv[(long)K * nvc + k] * dim + c];
        p->cell_meas[K] = shape == OR_BOX ? box_measure(dim, vv) : simplex_measure(dim, vv);
        for (int e = 0; e < fpc; e++) {
            double fv[12];
            int lv[4];
            const int t = face_vertices(dim, shape, e, lv);
            for (int k = 0; k < t; k++)
                for (int c = 0; c < dim; c++) fv[k * dim + c] = vv[lv[k] * dim + c];
            int E = p->cell_faces[(long)K * fpc + e];
            if (p->face_cells[2L * E] == K)
                p->face_meas[E] = (shape == OR_BOX && dim == 3) ? quad_face_measure(fv) : face_measure(dim, fv);
        }
    }
    const long C = nc, F = nf;
    p->Z_iF = calloc(n * F, sizeof(double));
    p->Z_iK = calloc(n * C, sizeof(double));
    p->N = calloc(n * n * C, sizeof(double));
    p->u = calloc(n * n * fpc * C, sizeof(double));
    p->m = calloc(n * C, sizeof(double));
    p->D = calloc(n * n * C, sizeof(double));
    p->w = calloc(n * fpc * C, sizeof(double));
    p->a = calloc(n * n * fpc * C, sizeof(double));
    p->r = calloc(n * n * C, sizeof(double));
    p->f = calloc(n * C, sizeof(double));
    p->v = calloc(n * fpc * C, sizeof(double));
    p->m_upw = calloc(n * F, sizeof(double));
    p->Z_upw = calloc(n * n * F, sizeof(double));
    p->b = calloc(n * n * p->S * C, sizeof(double));
    p->R_ijKe = calloc(n * n * fpc * C, sizeof(double));
    p->R_iK = calloc(n * C, sizeof(double));
    p->tags = calloc(n * F, 1);
    p->bc_values = calloc(n * F, sizeof(double));
    p->bc_dirichlet = calloc(n * F, sizeof(double));
    p->bc_mobility = calloc(n * F, sizeof(double));
    for (long k = 0; k < n * F; k++) p->m_upw[k] = 1.0;
    return p;
}

void or_destroy(or_problem *p)
{
    if (!p) return;
    void *ptrs[] = {p->pts, p->cv, p->cell_faces, p->face_cells, p->cell_meas, p->face_meas,
                    p->Z_iF, p->Z_iK, p->N, p->u, p->m, p->D, p->w, p->a, p->r, p->f, p->v,
                    p->m_upw, p->Z_upw, p->b, p->R_ijKe, p->R_iK, p->tags, p->bc_values,
                    p->bc_dirichlet, p->bc_mobility, p->rowptr, p->colidx, p->val, p->rhs};
    for (unsigned k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    free(p);
}

const char *or_error(const or_problem *p) { return p ? p->err : "null problem"; }

/* OpenMP threads of the parallel loops (the CPU baseline times 1 and many) */
void or_set_threads(int threads)
{
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : 1);
#else
    (void)threads;
#endif
}
int or_num_cells(const or_problem *p) { return p->nc; }
int or_num_faces(const or_problem *p) { return p->nf; }

void *or_array(or_problem *p, const char *name, long *size)
{
    const long C = p->nc, F = p->nf, n = p->n, fpc = p->fpc;
    struct { const char *nm; void *ptr; long sz; } tab[] = {
        {"Z_iF", p->Z_iF, n * F}, {"Z_iK", p->Z_iK, n * C}, {"N_ijK", p->N, n * n * C},
        {"u_ijKe", p->u, n * n * fpc * C}, {"m_iK", p->m, n * C}, {"D_ijK", p->D, n * n * C},
        {"w_iKe", p->w, n * fpc * C}, {"a_ijKe", p->a, n * n * fpc * C}, {"r_ijK", p->r, n * n * C},
        {"f_iK", p->f, n * C}, {"v_iKe", p->v, n * fpc * C}, {"m_iE_upw", p->m_upw, n * F},
        {"Z_ijE_upw", p->Z_upw, n * n * F}, {"b_ijK_storage", p->b, n * n * p->S * C},
        {"R_ijKe", p->R_ijKe, n * n * fpc * C}, {"R_iK", p->R_iK, n * C},
        {"cell_measure", p->cell_meas, C}, {"face_measure", p->face_meas, F},
        {"tags", p->tags, n * F}, {"bc_values", p->bc_values, n * F},
        {"bc_dirichlet", p->bc_dirichlet, n * F}, {"bc_mobility", p->bc_mobility, n * F},
        {"cell_faces", p->cell_faces, C * fpc}, {"face_cells", p->face_cells, 2 * F},
        {"rhs", p->rhs, p->rhs ? (long)p->nrows : 0},
    };
    for (unsigned k = 0; k < sizeof(tab) / sizeof(tab[0]); k++)
        if (!strcmp(tab[k].nm, name)) {
            if (size) *size = tab[k].sz;
            return tab[k].ptr;
        }
    if (size) *size = 0;
    return NULL;
}

static inline int local_index(const or_problem *p, int K, int E)
{
    /* getLocalIndex (mesh_helpers.h:253-265) */
    for (int e = 0; e < p->fpc; e++)
        if (p->cell_faces[(long)K * p->fpc + e] == E) return e;
    return 0;
}

/* Solver::setInitialCondition, face-average kernel (Solver.hpp:319-341) */
void or_init(or_problem *p)
{
    const int n = p->n;
#pragma omp parallel for
    for (int E = 0; E < p->nf; E++) {
        int K0 = p->face_cells[2L * E], K1 = p->face_cells[2L * E + 1];
        for (int i = 0; i < n; i++) {
            if (K1 < 0) ZF(i, E) = ZK(i, K0);
            else ZF(i, E) = 0.5 * (ZK(i, K0) + ZK(i, K1));
        }
    }
    memset(p->v, 0, sizeof(double) * (long)n * p->fpc * p->nc);
}

/* ------------------------------------------------------------------ */
/* Row lengths and CSR pattern: Solver.hpp:448-484, LinearSystem.h:26-34,
 * BoundaryConditions.hpp:283-299.  Row index E*n+i (BaseModel.h:259-263).
 * Column order: interior rows merge the two cells' face lists in ascending
 * face index (LinearSystem.h:108-246), boundary rows the sorted single-cell
 * list (BoundaryConditions.hpp:157-218); within a face, j ascending. */
static void sorted_faces(const or_problem *p, int K, int *perm)
{
    const int *fi = p->cell_faces + (long)K * p->fpc;
    for (int j = 0; j < p->fpc; j++) perm[j] = j;
    /* bubble sort of the local permutation (LinearSystem.h:129-136) */
    for (int k1 = p->fpc - 1; k1 > 0; k1--)
        for (int k2 = 0; k2 < k1; k2++)
            if (!(fi[perm[k2]] < fi[perm[k2 + 1]])) {
                int t = perm[k2]; perm[k2] = perm[k2 + 1]; perm[k2 + 1] = t;
            }
}

int or_setup_linear_system(or_problem *p)
{
    const int n = p->n, fpc = p->fpc;
    const long F = p->nf;
    p->nrows = (int)(n * F);
    free(p->rowptr); free(p->colidx); free(p->val); free(p->rhs);
    p->rowptr = malloc(sizeof(int) * (p->nrows + 1));
    p->rhs = calloc(p->nrows, sizeof(double));
    long nnz = 0;
    p->rowptr[0] = 0;
    for (long E = 0; E < F; E++)
        for (int i = 0; i < n; i++) {
            int len;
            if (p->face_cells[2 * E + 1] < 0)
                len = (p->tags[BCI(i, E)] == OR_FIXED_VALUE) ? 1 : fpc * n;
            else
                len = (2 * fpc - 1) * n;
            if (len <= 0) {
                snprintf(p->err, sizeof p->err, "MHFEM error: attempted to set invalid rowsLengths vector");
                return -1;
            }
            nnz += len;
            p->rowptr[E * n + i + 1] = (int)nnz;
        }
    p->nnz = nnz;
    p->colidx = malloc(sizeof(int) * nnz);
    p->val = calloc(nnz, sizeof(double));
#pragma omp parallel for
    for (long E = 0; E < F; E++) {
        int K0 = p->face_cells[2 * E], K1 = p->face_cells[2 * E + 1];
        for (int i = 0; i < n; i++) {
            int *c = p->colidx + p->rowptr[E * n + i];
            if (K1 < 0) {
                if (p->tags[BCI(i, E)] == OR_FIXED_VALUE) { c[0] = (int)(E * n + i); continue; }
                int perm[6];
                sorted_faces(p, K0, perm);
                int t = 0;
                for (int g = 0; g < fpc; g++)
                    for (int j = 0; j < n; j++) c[t++] = p->cell_faces[(long)K0 * fpc + perm[g]] * n + j;
            } else {
                int p0[6], p1[6], g0 = 0, g1 = 0, t = 0;
                const int *f0 = p->cell_faces + (long)K0 * fpc, *f1 = p->cell_faces + (long)K1 * fpc;
                sorted_faces(p, K0, p0);
                sorted_faces(p, K1, p1);
                while (g0 < fpc || g1 < fpc) {
                    int F0 = g0 < fpc ? f0[p0[g0]] : 0x7fffffff, F1 = g1 < fpc ? f1[p1[g1]] : 0x7fffffff;
                    int Fm = F0 < F1 ? F0 : F1;
                    for (int j = 0; j < n; j++) c[t++] = Fm * n + j;
                    if (F0 == Fm) g0++;
                    if (F1 == Fm) g1++;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* SecondaryCoefficients (SecondaryCoefficients.h) */
static double A_no_adv(const or_problem *p, int i, int j, int K, int e, int f)
{
    /* A_ijKEF_no_advection (SecondaryCoefficients.h:189-211) */
    double SQinvR = 0;
    for (int x = 0; x < p->n; x++) SQinvR += b_e(p, i, x, K, e) * p->R_ijKe[IJKE(x, j, K, f)];
    return b_ef(p, i, j, K, e, f) - SQinvR;
}

static double A_adv(const or_problem *p, int i, int j, int K, int E, int e, int F, int f)
{
    /* A_ijKEF_advection (SecondaryCoefficients.h:215-259) */
    const double mu = MUPW(i, E);
    double SQinvR = 0;
    for (int x = 0; x < p->n; x++) {
        double S = mu * b_e(p, i, x, K, e);
        if (p->advection == OR_IMPLICIT_UPWIND) {
            double vel = p->a[IJKE(i, x, K, e)] + p->u[IJKE(i, x, K, e)];
            if (vel >= 0) S += vel; else S -= vel;
        }
        SQinvR += S * p->R_ijKe[IJKE(x, j, K, f)];
    }
    double T = mu * b_ef(p, i, j, K, e, f);
    if (p->advection == OR_IMPLICIT_UPWIND && E == F) {
        double vel = p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)];
        if (vel < 0) T -= 2 * vel;
    }
    if (p->advection == OR_IMPLICIT_TRACE && E == F) {
        double vel = p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)];
        T -= vel;
    }
    return T - SQinvR;
}

static double A_interior(const or_problem *p, int i, int j, int K, int E, int e, int F, int f)
{
    /* A_ijKEF (SecondaryCoefficients.h:261-278) */
    if (p->advection == OR_EXPLICIT_UPWIND) return A_no_adv(p, i, j, K, e, f);
    return A_adv(p, i, j, K, E, e, F, f);
}

static double RHS_no_adv(const or_problem *p, int i, int K, int e)
{
    /* RHS_iKE_no_advection (SecondaryCoefficients.h:282-297) */
    double value = p->w[IKE(i, K, e)];
    for (int j = 0; j < p->n; j++) value += b_e(p, i, j, K, e) * p->R_iK[IK(j, K)];
    return value;
}

static double RHS_adv(const or_problem *p, int i, int K, int E, int e)
{
    /* RHS_iKE_advection (SecondaryCoefficients.h:301-333) */
    const double mu = MUPW(i, E);
    double value = mu * p->w[IKE(i, K, e)];
    for (int j = 0; j < p->n; j++) {
        double S = mu * b_e(p, i, j, K, e);
        const double vel = p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)];
        if (p->advection == OR_IMPLICIT_UPWIND) {
            if (vel >= 0) S += vel; else S -= vel;
        }
        value += S * p->R_iK[IK(j, K)];
        if (p->advection == OR_EXPLICIT_UPWIND) value += ZUPW(i, j, E) * vel;
    }
    return value;
}

static double RHS_interior(const or_problem *p, int i, int K, int E, int e)
{
    /* RHS_iKE (SecondaryCoefficients.h:337-349) */
    if (p->advection == OR_EXPLICIT_UPWIND) return RHS_no_adv(p, i, K, e);
    return RHS_adv(p, i, K, E, e);
}

/* ------------------------------------------------------------------ */
/* Solver::preIterate (Solver.hpp:600-892) */
void or_pre_iterate(or_problem *p, double time, double tau)
{
    (void)time;
    const int n = p->n, fpc = p->fpc, dim = p->dim;
    const int C = p->nc, F = p->nf;
    /* K2: MassMatrix::update over (K,i,j) (Solver.hpp:627-639) */
#pragma omp parallel for
    for (int K = 0; K < C; K++) {
        double P[12], out[14];
        if (p->shape == OR_BOX) {
            /* MassMatrix.h:46-54, 107-115 (edge: h_x = |K|), :160-176, 231-247
             * (quadrangle: h from vertices 0 and 2, an axis-aligned rectangle) */
            const int *cvK = p->cv + (long)K * p->nvc;
            double hx, hy = 0, hz = 0;
            if (dim == 1) hx = p->cell_meas[K];
            else {
                /* the opposite corner: vertex 2 of a quadrangle, 6 of a hexahedron
                 * (MassMatrix.h:168-172, 298-304) */
                const long a = cvK[0], c = cvK[dim == 2 ? 2 : 6];
                hx = p->pts[c * dim] - p->pts[a * dim];
                hy = p->pts[c * dim + 1] - p->pts[a * dim + 1];
                if (dim == 3) hz = p->pts[c * dim + 2] - p->pts[a * dim + 2];
            }
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) {
                    const double D = p->D[IJK(i, j, K)];
                    if (dim == 1) BST(i, j, K, 0) = 2 * D / hx;
                    else if (dim == 2) {
                        BST(i, j, K, 0) = 2 * D * hx / hy;
                        BST(i, j, K, 1) = 2 * D * hy / hx;
                    } else { /* MassMatrix.h:306-312 */
                        BST(i, j, K, 0) = 2 * D * hy * hz / hx;
                        BST(i, j, K, 1) = 2 * D * hx * hz / hy;
                        BST(i, j, K, 2) = 2 * D * hx * hy / hz;
                    }
                }
            continue;
        }
        for (int k = 0; k < fpc; k++)
            for (int c = 0; c < dim; c++) P[k * dim + c] = p->pts[(long)p->cv[(long)K * fpc + k] * dim + c];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                or_local_mass_inverse(dim, p->lumping, P, p->cell_meas[K], p->D[IJK(i, j, K)], out);
                for (int s = 0; s < p->S; s++) BST(i, j, K, s) = out[s];
            }
    }
    /* K4: upwinded mobility over (E,i) (Solver.hpp:663-727) */
    if (p->do_mob) {
#pragma omp parallel for
        for (int E = 0; E < F; E++) {
            const int K1 = p->face_cells[2L * E], K2 = p->face_cells[2L * E + 1];
            const int e1 = local_index(p, K1, E);
            for (int i = 0; i < n; i++) {
                double mu;
                if (K2 < 0) {
                    int inflow = 0;
                    for (int j = 0; j < n; j++)
                        if (p->v[IKE(j, K1, e1)] < 0) { inflow = 1; break; }
                    mu = inflow ? p->bc_mobility[BCI(i, E)] : p->m[IK(i, K1)];
                } else {
                    const int e2 = local_index(p, K2, E);
                    const double vel = p->v[IKE(i, K1, e1)] - p->v[IKE(i, K2, e2)];
                    if (vel == 0) mu = 0.5 * (p->m[IK(i, K1)] + p->m[IK(i, K2)]);
                    else if (vel > 0) mu = p->m[IK(i, K1)];
                    else mu = p->m[IK(i, K2)];
                }
                MUPW(i, E) = mu;
            }
        }
    }
    /* K5: explicit upwind of Z over (E,j,i) (Solver.hpp:729-769) */
    if (p->advection == OR_EXPLICIT_UPWIND) {
#pragma omp parallel for
        for (int E = 0; E < F; E++) {
            const int K1 = p->face_cells[2L * E], K2 = p->face_cells[2L * E + 1];
            const int e1 = local_index(p, K1, E);
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    const double apu = p->a[IJKE(i, j, K1, e1)] + p->u[IJKE(i, j, K1, e1)];
                    double z;
                    if (apu > 0.0) z = ZK(j, K1);
                    else if (apu == 0.0) z = 0;
                    else if (K2 >= 0) z = ZK(j, K2);
                    else z = ZF(j, E);
                    ZUPW(i, j, E) = z;
                }
        }
    }
    /* K6: R_ijKe and R_iK over (K,i) (Solver.hpp:793-815) */
#pragma omp parallel for
    for (int K = 0; K < C; K++) {
        const int *fi = p->cell_faces + (long)K * fpc;
        const double mK = p->cell_meas[K];
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++)
                for (int e = 0; e < fpc; e++) {
                    /* R_ijKe (SecondaryCoefficients.h:50-86) */
                    double R = 0;
                    for (int f = 0; f < fpc; f++) R += MUPW(i, fi[f]) * b_ef(p, i, j, K, f, e);
                    if (p->advection == OR_IMPLICIT_UPWIND) {
                        double vel = p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)];
                        if (vel < 0) R -= 2 * vel;
                    }
                    if (p->advection == OR_IMPLICIT_TRACE)
                        R -= p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)];
                    p->R_ijKe[IJKE(i, j, K, e)] = R;
                }
            /* R_iK (SecondaryCoefficients.h:88-127) */
            double R = 0;
            for (int j = 0; j < n; j++) R += p->N[IJK(i, j, K)] * ZK(j, K);
            R *= mK / tau;
            R += mK * p->f[IK(i, K)];
            for (int e = 0; e < fpc; e++) R -= MUPW(i, fi[e]) * p->w[IKE(i, K, e)];
            if (p->advection == OR_EXPLICIT_UPWIND) {
                double aux = 0;
                for (int j = 0; j < n; j++)
                    for (int e = 0; e < fpc; e++)
                        aux += (p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)]) * ZUPW(i, j, fi[e]);
                R -= aux;
            }
            p->R_iK[IK(i, K)] = R;
        }
    }
    /* K7: Q, singular-row fix, LU and in-place Q^{-1} solves (Solver.hpp:817-871) */
#pragma omp parallel for
    for (int K = 0; K < C; K++) {
        const int *fi = p->cell_faces + (long)K * fpc;
        const double mK = p->cell_meas[K];
        double Q[64], rhs[8];
        for (int i = 0; i < n; i++) {
            int singular = 1;
            for (int j = 0; j < n; j++) {
                /* Q_ijK (SecondaryCoefficients.h:129-159) */
                double q = 0;
                for (int e = 0; e < fpc; e++) {
                    q += MUPW(i, fi[e]) * b_e(p, i, j, K, e) - p->u[IJKE(i, j, K, e)];
                    if (p->advection == OR_IMPLICIT_UPWIND) {
                        double vel = p->a[IJKE(i, j, K, e)] + p->u[IJKE(i, j, K, e)];
                        if (vel >= 0) q += vel; else q -= vel;
                    }
                }
                q += mK / tau * p->N[IJK(i, j, K)] + mK * p->r[IJK(i, j, K)];
                Q[i * n + j] = q;
                if (q != 0.0) singular = 0;
            }
            if (singular) {
                Q[i * n + i] = 1.0;
                p->R_iK[IK(i, K)] += ZK(i, K);
            }
        }
        lu_factorize(Q, n);
        for (int i = 0; i < n; i++) rhs[i] = p->R_iK[IK(i, K)];
        lu_solve(Q, n, rhs);
        for (int i = 0; i < n; i++) p->R_iK[IK(i, K)] = rhs[i];
        for (int j = 0; j < n; j++)
            for (int e = 0; e < fpc; e++) {
                for (int i = 0; i < n; i++) rhs[i] = p->R_ijKe[IJKE(i, j, K, e)];
                lu_solve(Q, n, rhs);
                for (int i = 0; i < n; i++) p->R_ijKe[IJKE(i, j, K, e)] = rhs[i];
            }
    }
}

/* ------------------------------------------------------------------ */
/* Solver::assembleLinearSystem (Solver.hpp:933-959) */
static void boundary_row(or_problem *p, int E, int i, int row)
{
    /* BoundaryConditions::setMatrixElements (BoundaryConditions.hpp:335-508) */
    const int n = p->n, fpc = p->fpc;
    const int tag = p->tags[BCI(i, E)];
    double *vals = p->val + p->rowptr[row];
    if (tag == OR_FIXED_VALUE) {
        vals[0] = 1.0;
        p->rhs[row] = p->bc_dirichlet[BCI(i, E)];
        return;
    }
    const int K = p->face_cells[2L * E];
    const int *fi = p->cell_faces + (long)K * fpc;
    const int e = local_index(p, K, E);
    int perm[6];
    sorted_faces(p, K, perm);
    /* which coefficient set (BoundaryConditions.hpp:11-130) */
    int inflow_identity = 0, use_adv = 0;
    if (tag == OR_FIXED_FLUX) use_adv = 1;
    if (tag == OR_ADVECTIVE_OUTFLOW) {
        const double vel = p->a[IJKE(i, i, K, e)] + p->u[IJKE(i, i, K, e)];
        if (vel < 0) inflow_identity = 1;
    }
#define BA(jj, FF, ff) (inflow_identity ? ((i == (jj) && E == (FF)) ? 1.0 : 0.0) \
                        : (use_adv ? A_adv(p, i, jj, K, E, e, FF, ff) : A_no_adv(p, i, jj, K, e, ff)))
    /* RowSetter::setRow (BoundaryConditions.hpp:132-222) */
    const double diag = BA(i, E, e);
    int t = 0;
    for (int g = 0; g < fpc; g++) {
        const int f = perm[g];
        for (int j = 0; j < n; j++) {
            if (j == i && fi[f] == E) vals[t++] = 1.0;
            else vals[t++] = BA(j, fi[f], f) / diag;
        }
    }
#undef BA
    double bval;
    if (tag == OR_ADVECTIVE_OUTFLOW) {
        bval = 0;
        bval += inflow_identity ? ZK(i, K) : RHS_no_adv(p, i, K, e);
    } else {
        bval = -p->bc_values[BCI(i, E)] * p->face_meas[E];
        bval += use_adv ? RHS_adv(p, i, K, E, e) : RHS_no_adv(p, i, K, e);
    }
    p->rhs[row] = bval / diag;
}

static void interior_row(or_problem *p, int E, int i, int row)
{
    /* LinearSystem::setMatrixElements (LinearSystem.h:66-256) and
     * RightHandSide::getValue (RightHandSide.h:18-45) */
    const int n = p->n, fpc = p->fpc;
    const int K0 = p->face_cells[2L * E], K1 = p->face_cells[2L * E + 1];
    const int *f0 = p->cell_faces + (long)K0 * fpc, *f1 = p->cell_faces + (long)K1 * fpc;
    int p0[6], p1[6];
    sorted_faces(p, K0, p0);
    sorted_faces(p, K1, p1);
    const int e0 = local_index(p, K0, E), e1 = local_index(p, K1, E);
    const double d = A_interior(p, i, i, K0, E, e0, E, e0) + A_interior(p, i, i, K1, E, e1, E, e1);
    double *vals = p->val + p->rowptr[row];
    int t = 0, g0 = 0, g1 = 0;
    while (g0 < fpc && g1 < fpc) {
        const int a0 = p0[g0], a1 = p1[g1];
        if (f0[a0] < f1[a1]) {
            for (int j = 0; j < n; j++) vals[t++] = A_interior(p, i, j, K0, E, e0, f0[a0], a0) / d;
            g0++;
        } else if (f0[a0] == f1[a1]) {
            for (int j = 0; j < n; j++) {
                if (j == i) vals[t++] = 1.0;
                else
                    vals[t++] = (A_interior(p, i, j, K0, E, e0, f0[a0], a0) +
                                 A_interior(p, i, j, K1, E, e1, f1[a1], a1)) / d;
            }
            g0++;
            g1++;
        } else {
            for (int j = 0; j < n; j++) vals[t++] = A_interior(p, i, j, K1, E, e1, f1[a1], a1) / d;
            g1++;
        }
    }
    while (g0 < fpc) {
        const int a0 = p0[g0++];
        for (int j = 0; j < n; j++) vals[t++] = A_interior(p, i, j, K0, E, e0, f0[a0], a0) / d;
    }
    while (g1 < fpc) {
        const int a1 = p1[g1++];
        for (int j = 0; j < n; j++) vals[t++] = A_interior(p, i, j, K1, E, e1, f1[a1], a1) / d;
    }
    double rhs = 0.0;
    rhs += RHS_interior(p, i, K0, E, e0);
    rhs += RHS_interior(p, i, K1, E, e1);
    p->rhs[row] = rhs / d;
}

void or_assemble(or_problem *p, double time, double tau)
{
    (void)time; (void)tau;
    const int n = p->n;
#pragma omp parallel for
    for (int E = 0; E < p->nf; E++)
        for (int i = 0; i < n; i++) {
            const int row = E * n + i;
            if (p->face_cells[2L * E + 1] < 0) boundary_row(p, E, i, row);
            else interior_row(p, E, i, row);
        }
}

/* ------------------------------------------------------------------ */
/* Solver::postIterate (Solver.hpp:1273-1324) */
void or_post_iterate(or_problem *p, double time, double tau)
{
    (void)time; (void)tau;
    const int n = p->n, fpc = p->fpc;
#pragma omp parallel for
    for (int K = 0; K < p->nc; K++) {
        const int *fi = p->cell_faces + (long)K * fpc;
        /* K9: Z_iK (SecondaryCoefficients.h:163-185) */
        for (int i = 0; i < n; i++) {
            double Z = 0;
            for (int f = 0; f < fpc; f++)
                for (int j = 0; j < n; j++) Z += p->R_ijKe[IJKE(i, j, K, f)] * ZF(j, fi[f]);
            Z += p->R_iK[IK(i, K)];
            ZK(i, K) = Z;
        }
        /* K10: v_iKE (SecondaryCoefficients.h:16-48), uses the new Z_iK */
        if (p->do_mob) {
            for (int i = 0; i < n; i++)
                for (int e = 0; e < fpc; e++) {
                    double sK = 0, sE = 0;
                    for (int j = 0; j < n; j++) {
                        const double bb = b_e(p, i, j, K, e);
                        sK += bb * ZK(j, K);
                        for (int f = 0; f < fpc; f++) sE += b_ef(p, i, j, K, e, f) * ZF(j, fi[f]);
                    }
                    p->v[IKE(i, K, e)] = sK - sE + p->w[IKE(i, K, e)];
                }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Sparse kernels and Krylov solvers (restated published methods). */
void or_spmv(int nrows, const int *rp, const int *col, const double *val, const double *x, double *y)
{
#pragma omp parallel for
    for (int r = 0; r < nrows; r++) {
        double s = 0;
        for (int k = rp[r]; k < rp[r + 1]; k++) s += val[k] * x[col[k]];
        y[r] = s;
    }
}

/* ILU(0), IKJ variant: Saad, Iterative Methods for Sparse Linear Systems,
 * 2nd ed., Algorithm 10.4.  Pinned in tests by (LU)_ij == A_ij on the pattern. */
int or_ilu0_factorize(int nrows, const int *rp, const int *col, double *a)
{
    int *iw = malloc(sizeof(int) * nrows);
    int *dp = malloc(sizeof(int) * nrows);
    for (int r = 0; r < nrows; r++) iw[r] = -1;
    int rc = 0;
    for (int r = 0; r < nrows && !rc; r++) {
        dp[r] = -1;
        for (int k = rp[r]; k < rp[r + 1]; k++) {
            iw[col[k]] = k;
            if (col[k] == r) dp[r] = k;
        }
        if (dp[r] < 0) { rc = -1; break; }
        for (int k = rp[r]; k < rp[r + 1] && col[k] < r; k++) {
            const int c = col[k];
            const double piv = a[k] / a[dp[c]];
            a[k] = piv;
            for (int kk = dp[c] + 1; kk < rp[c + 1]; kk++) {
                const int pos = iw[col[kk]];
                if (pos >= 0) a[pos] -= piv * a[kk];
            }
        }
        if (a[dp[r]] == 0.0) rc = -1;
        for (int k = rp[r]; k < rp[r + 1]; k++) iw[col[k]] = -1;
    }
    free(iw);
    free(dp);
    return rc;
}

void or_ilu0_apply(int nrows, const int *rp, const int *col, const double *lu, const double *x,
                   double *y)
{
    /* forward: unit lower L */
    for (int r = 0; r < nrows; r++) {
        double s = x[r];
        for (int k = rp[r]; k < rp[r + 1] && col[k] < r; k++) s -= lu[k] * y[col[k]];
        y[r] = s;
    }
    /* backward: U with diagonal */
    for (int r = nrows - 1; r >= 0; r--) {
        double s = y[r], dg = 1.0;
        for (int k = rp[r]; k < rp[r + 1]; k++) {
            if (col[k] > r) s -= lu[k] * y[col[k]];
            else if (col[k] == r) dg = lu[k];
        }
        y[r] = s / dg;
    }
}

/* ILUT(p, tau) (Saad, Iterative Methods, Alg. 10.6; TNL's "ilut"
 * preconditioner): row by row, the row of A in a dense work vector w is
 * eliminated against the finished rows of U in increasing column order,
 * fill-in included; a multiplier w_k / u_kk below tau_i = tau ||a_i||_2 is
 * dropped; then entries below tau_i are dropped and only the p largest in
 * magnitude of the L part and of the U part are kept (ties: smaller column),
 * the diagonal always (a zero pivot becomes tau_i, or 1).  Output: CSR with
 * sorted columns, L (unit diagonal not stored) left of the diagonal entry of
 * U.  Returns nnz, or -1 on allocation failure; *rp_out etc. malloc'ed. */
typedef struct { int j; double v; } ilut_entry;
static int ilut_by_mag(const void *a, const void *b)
{
    const ilut_entry *x = a, *y = b;
    const double ax = fabs(x->v), ay = fabs(y->v);
    if (ax != ay) return ax > ay ? -1 : 1;
    return x->j - y->j;
}
static int ilut_by_col(const void *a, const void *b)
{
    return ((const ilut_entry *)a)->j - ((const ilut_entry *)b)->j;
}
long or_ilut_factorize(int n, const int *rp, const int *col, const double *val, int p, double tau, int **rp_out,
                       int **col_out, double **val_out)
{
    long cap = (long)rp[n] * 2 + n, nnz = 0;
    int *frp = malloc(sizeof(int) * (n + 1)), *fcol = malloc(sizeof(int) * cap), *fdg = malloc(sizeof(int) * n);
    double *fval = malloc(sizeof(double) * cap), *w = calloc(n, sizeof(double));
    int *mark = malloc(sizeof(int) * n), *idx = malloc(sizeof(int) * n), *done = calloc(n, sizeof(int));
    ilut_entry *ent = malloc(sizeof(ilut_entry) * n);
    for (int i = 0; i < n; i++) mark[i] = -1;
    frp[0] = 0;
    for (int i = 0; i < n; i++) {
        int cnt = 0;
        double nrm = 0;
        for (int q = rp[i]; q < rp[i + 1]; q++) {
            w[col[q]] = val[q];
            mark[col[q]] = i;
            idx[cnt++] = col[q];
            nrm += val[q] * val[q];
        }
        if (mark[i] != i) { w[i] = 0; mark[i] = i; idx[cnt++] = i; }
        const double ti = tau * sqrt(nrm);
        for (;;) {
            /* the smallest unprocessed column left of the diagonal */
            int k = -1;
            for (int t = 0; t < cnt; t++)
                if (idx[t] < i && done[idx[t]] != i + 1 && (k < 0 || idx[t] < k)) k = idx[t];
            if (k < 0) break;
            done[k] = i + 1;
            if (w[k] == 0.0) continue;
            w[k] /= fval[fdg[k]];
            if (fabs(w[k]) < ti) { w[k] = 0.0; continue; }
            for (long q = fdg[k] + 1; q < frp[k + 1]; q++) {
                const int j = fcol[q];
                if (mark[j] != i) { mark[j] = i; w[j] = 0.0; idx[cnt++] = j; }
                w[j] -= w[k] * fval[q];
            }
        }
        int nl = 0, nu = 0;
        ilut_entry *L = ent, *U;
        for (int t = 0; t < cnt; t++) {
            const int j = idx[t];
            if (j < i && w[j] != 0.0 && fabs(w[j]) >= ti) L[nl++] = (ilut_entry){j, w[j]};
        }
        U = ent + nl;
        for (int t = 0; t < cnt; t++) {
            const int j = idx[t];
            if (j > i && w[j] != 0.0 && fabs(w[j]) >= ti) U[nu++] = (ilut_entry){j, w[j]};
        }
        qsort(L, nl, sizeof(ilut_entry), ilut_by_mag);
        qsort(U, nu, sizeof(ilut_entry), ilut_by_mag);
        if (nl > p) nl = p;
        if (nu > p) nu = p;
        qsort(L, nl, sizeof(ilut_entry), ilut_by_col);
        qsort(U, nu, sizeof(ilut_entry), ilut_by_col);
        if (nnz + nl + nu + 1 > cap) {
            cap = 2 * cap + nl + nu + 1;
            fcol = realloc(fcol, sizeof(int) * cap);
            fval = realloc(fval, sizeof(double) * cap);
        }
        for (int t = 0; t < nl; t++) { fcol[nnz] = L[t].j; fval[nnz++] = L[t].v; }
        double d = w[i];
        if (d == 0.0) d = ti > 0 ? ti : 1.0;
        fdg[i] = (int)nnz;
        fcol[nnz] = i;
        fval[nnz++] = d;
        for (int t = 0; t < nu; t++) { fcol[nnz] = U[t].j; fval[nnz++] = U[t].v; }
        frp[i + 1] = (int)nnz;
        for (int t = 0; t < cnt; t++) w[idx[t]] = 0.0;
    }
    free(w); free(mark); free(idx); free(done); free(ent); free(fdg);
    *rp_out = frp;
    *col_out = fcol;
    *val_out = fval;
    return nnz;
}

static int ilut_p = 10;
static double ilut_tau = 1e-4;
void or_set_ilut(int p, double tau)
{
    ilut_p = p;
    ilut_tau = tau;
}

typedef struct {
    int n;
    const int *rp, *col;
    const double *val;
    int prec;
    double *dinv, *lu, *tmp;
    int *frp, *fcol;  /* ILUT: the factors' own pattern */
} krylov_op;

static void prec_apply(krylov_op *op, const double *x, double *y)
{
    const int n = op->n;
    if (op->prec == OR_PREC_DIAGONAL) {
#pragma omp parallel for
        for (int r = 0; r < n; r++) y[r] = x[r] * op->dinv[r];
    } else if (op->prec == OR_PREC_ILU0) {
        or_ilu0_apply(n, op->rp, op->col, op->lu, x, y);
    } else if (op->prec == OR_PREC_ILUT) {
        or_ilu0_apply(n, op->frp, op->fcol, op->lu, x, y);
    } else {
        memcpy(y, x, sizeof(double) * n);
    }
}
/* y = M^{-1} A x */
static void pmatvec(krylov_op *op, const double *x, double *y)
{
    or_spmv(op->n, op->rp, op->col, op->val, x, op->tmp);
    prec_apply(op, op->tmp, y);
}
/* r = M^{-1}(b - A x) */
static void presidue(krylov_op *op, const double *x, const double *b, double *r)
{
    or_spmv(op->n, op->rp, op->col, op->val, x, op->tmp);
#pragma omp parallel for
    for (int k = 0; k < op->n; k++) op->tmp[k] = b[k] - op->tmp[k];
    prec_apply(op, op->tmp, r);
}
static double dot(int n, const double *x, const double *y)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s)
    for (int k = 0; k < n; k++) s += x[k] * y[k];
    return s;
}

/* BiCGStab, left preconditioned (van der Vorst 1992), with the restart on
 * rho breakdown; stopping test ||r_k|| / ||M^{-1} b|| < tol.  The two other
 * textbook breakdowns, (r*, A p) = 0 and omega = 0, also restart from the
 * current x (where the plain recurrence would continue with inf/NaN scalars). */
static int bicgstab(krylov_op *op, const double *b, double *x, double tol, int maxit, int *iters,
                    double *res)
{
    const int n = op->n;
    double *r = malloc(sizeof(double) * n), *r_ast = malloc(sizeof(double) * n),
           *pp = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n),
           *s = malloc(sizeof(double) * n), *As = malloc(sizeof(double) * n);
    prec_apply(op, b, s);
    double b_norm = sqrt(dot(n, s, s));
    if (b_norm == 0.0) b_norm = 1.0;
    presidue(op, x, b, r);
    memcpy(pp, r, sizeof(double) * n);
    memcpy(r_ast, r, sizeof(double) * n);
    double rho = dot(n, r, r_ast), r_ast_sqnorm = rho;
    const double eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
    double resid = sqrt(rho) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && !isnan(resid)) {
        pmatvec(op, pp, Ap);
        const double den = dot(n, Ap, r_ast), alpha = rho / den;
        if (!(fabs(den) > 0.0) || !isfinite(alpha)) {
            /* breakdown (r*, A p) = 0: restart from the current x */
            presidue(op, x, b, r);
            memcpy(pp, r, sizeof(double) * n);
            memcpy(r_ast, r, sizeof(double) * n);
            rho = r_ast_sqnorm = dot(n, r, r_ast);
            it++;
            resid = sqrt(dot(n, r, r)) / b_norm;
            continue;
        }
#pragma omp parallel for
        for (int k = 0; k < n; k++) s[k] = r[k] - alpha * Ap[k];
        pmatvec(op, s, As);
        double As_s = 0, As_As = 0;
#pragma omp parallel for reduction(+ : As_s, As_As)
        for (int k = 0; k < n; k++) { As_s += As[k] * s[k]; As_As += As[k] * As[k]; }
        const double omega = As_As > 0 ? As_s / As_As : 0.0;
#pragma omp parallel for
        for (int k = 0; k < n; k++) {
            x[k] += alpha * pp[k] + omega * s[k];
            r[k] = s[k] - omega * As[k];
        }
        const double rho_old = rho;
        rho = dot(n, r, r_ast);
        if (!(fabs(rho) >= eps2 * r_ast_sqnorm) || omega == 0.0) {
            presidue(op, x, b, r);
            memcpy(pp, r, sizeof(double) * n);
            memcpy(r_ast, r, sizeof(double) * n);
            rho = r_ast_sqnorm = dot(n, r, r_ast);
        } else {
            const double beta = (rho / rho_old) * (alpha / omega);
#pragma omp parallel for
            for (int k = 0; k < n; k++) pp[k] = r[k] + beta * (pp[k] - omega * Ap[k]);
        }
        it++;
        resid = sqrt(dot(n, r, r)) / b_norm;
    }
    *iters = it;
    *res = resid;
    free(r); free(r_ast); free(pp); free(Ap); free(s); free(As);
    return resid < tol;
}

/* Restarted GMRES(m), left preconditioned (Saad & Schultz 1986), classical
 * Gram-Schmidt with one re-orthogonalisation pass, Givens rotations. */
static int gmres(krylov_op *op, const double *b, double *x, double tol, int maxit, int m, int *iters,
                 double *res)
{
    const int n = op->n;
    if (m < 1) m = 30;
    double *V = malloc(sizeof(double) * (long)n * (m + 1));
    double *w = malloc(sizeof(double) * n);
    double *H = calloc((long)(m + 1) * m, sizeof(double));
    double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
    double *g = malloc(sizeof(double) * (m + 1)), *y = malloc(sizeof(double) * m);
    double *h = malloc(sizeof(double) * (m + 1)), *h2 = malloc(sizeof(double) * (m + 1));
    prec_apply(op, b, w);
    double b_norm = sqrt(dot(n, w, w));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    double resid;
    presidue(op, x, b, V);
    double beta = sqrt(dot(n, V, V));
    resid = beta / b_norm;
    while (resid >= tol && it < maxit && !isnan(resid)) {
        for (int k = 0; k < n; k++) V[k] /= beta;
        for (int k = 0; k <= m; k++) g[k] = 0;
        g[0] = beta;
        int j;
        for (j = 0; j < m && it < maxit; j++) {
            pmatvec(op, V + (long)j * n, w);
            for (int pass = 0; pass < 2; pass++) {
                for (int k = 0; k <= j; k++) h2[k] = dot(n, V + (long)k * n, w);
#pragma omp parallel for
                for (int q = 0; q < n; q++) {
                    double s = w[q];
                    for (int k = 0; k <= j; k++) s -= h2[k] * V[(long)k * n + q];
                    w[q] = s;
                }
                for (int k = 0; k <= j; k++) h[k] = pass ? h[k] + h2[k] : h2[k];
            }
            h[j + 1] = sqrt(dot(n, w, w));
            if (h[j + 1] != 0.0)
                for (int q = 0; q < n; q++) V[(long)(j + 1) * n + q] = w[q] / h[j + 1];
            for (int k = 0; k < j; k++) {
                double t = cs[k] * h[k] + sn[k] * h[k + 1];
                h[k + 1] = -sn[k] * h[k] + cs[k] * h[k + 1];
                h[k] = t;
            }
            double den = sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
            cs[j] = den > 0 ? h[j] / den : 1.0;
            sn[j] = den > 0 ? h[j + 1] / den : 0.0;
            h[j] = cs[j] * h[j] + sn[j] * h[j + 1];
            h[j + 1] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int k = 0; k <= j; k++) H[(long)k * m + j] = h[k];
            it++;
            resid = fabs(g[j + 1]) / b_norm;
            if (resid < tol) { j++; break; }
        }
        /* back substitution and update */
        for (int k = j - 1; k >= 0; k--) {
            double s = g[k];
            for (int q = k + 1; q < j; q++) s -= H[(long)k * m + q] * y[q];
            y[k] = s / H[(long)k * m + k];
        }
#pragma omp parallel for
        for (int q = 0; q < n; q++) {
            double s = 0;
            for (int k = 0; k < j; k++) s += y[k] * V[(long)k * n + q];
            x[q] += s;
        }
        presidue(op, x, b, V);
        beta = sqrt(dot(n, V, V));
        resid = beta / b_norm;
    }
    *iters = it;
    *res = resid;
    free(V); free(w); free(H); free(cs); free(sn); free(g); free(y); free(h); free(h2);
    return resid < tol;
}

/* Preconditioned CG (Hestenes & Stiefel 1952): z = M^{-1} r, stopping test
 * ||z|| / ||M^{-1} b|| < tol (the preconditioned residual, as for the other
 * solvers).  Meant for symmetric positive definite systems. */
static int cg(krylov_op *op, const double *b, double *x, double tol, int maxit, int *iters, double *res)
{
    const int n = op->n;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *pp = malloc(sizeof(double) * n),
           *q = malloc(sizeof(double) * n);
    prec_apply(op, b, z);
    double b_norm = sqrt(dot(n, z, z));
    if (b_norm == 0.0) b_norm = 1.0;
    or_spmv(n, op->rp, op->col, op->val, x, r);
#pragma omp parallel for
    for (int k = 0; k < n; k++) r[k] = b[k] - r[k];
    prec_apply(op, r, z);
    memcpy(pp, z, sizeof(double) * n);
    double rho = dot(n, r, z), resid = sqrt(dot(n, z, z)) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && isfinite(resid)) {
        or_spmv(n, op->rp, op->col, op->val, pp, q);
        const double alpha = rho / dot(n, pp, q);
#pragma omp parallel for
        for (int k = 0; k < n; k++) {
            x[k] += alpha * pp[k];
            r[k] -= alpha * q[k];
        }
        prec_apply(op, r, z);
        const double rho_new = dot(n, r, z);
        resid = sqrt(dot(n, z, z)) / b_norm;
        const double beta = rho_new / rho;
        rho = rho_new;
#pragma omp parallel for
        for (int k = 0; k < n; k++) pp[k] = z[k] + beta * pp[k];
        it++;
    }
    *iters = it;
    *res = resid;
    free(r); free(z); free(pp); free(q);
    return resid < tol;
}

/* TFQMR (Freund 1993) on the left-preconditioned system M^{-1} A x = M^{-1} b,
 * in the two-half-step form (Kelley 1995, Alg. tfqmr), shadow vector r_0.
 * An iteration is one pair of half steps (two operator applications).  The
 * quasi-residual bound tau sqrt(m + 1) decides when to look at the true
 * preconditioned residual, which must then be below tol (else the iteration
 * goes on).  On a breakdown ((r_0, v) = 0 or rho = 0, which the MHFEM systems
 * with their identity Dirichlet rows do reach) it restarts from the current
 * x, as BiCGStab does here. */
static int tfqmr(krylov_op *op, const double *b, double *x, double tol, int maxit, int *iters, double *res)
{
    const int n = op->n;
    double *rt = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n), *y1 = malloc(sizeof(double) * n),
           *y2 = malloc(sizeof(double) * n), *u1 = malloc(sizeof(double) * n), *u2 = malloc(sizeof(double) * n),
           *v = malloc(sizeof(double) * n), *d = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n);
    prec_apply(op, b, r);
    double b_norm = sqrt(dot(n, r, r));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0, converged = 0;
    double tau, theta, eta, rho, resid;
restart:
    presidue(op, x, b, rt);
    memcpy(w, rt, sizeof(double) * n);
    memcpy(y1, rt, sizeof(double) * n);
    memset(d, 0, sizeof(double) * n);
    pmatvec(op, y1, v);
    memcpy(u1, v, sizeof(double) * n);
    tau = sqrt(dot(n, rt, rt));
    theta = eta = 0;
    rho = tau * tau;
    resid = tau / b_norm;
    if (resid < tol) converged = 1;
    while (!converged && it < maxit && isfinite(resid)) {
        it++;
        const double sigma = dot(n, rt, v), alpha = rho / sigma;
        if (!(fabs(sigma) > 0.0) || !isfinite(alpha)) goto restart;
#pragma omp parallel for
        for (int k = 0; k < n; k++) y2[k] = y1[k] - alpha * v[k];
        pmatvec(op, y2, u2);
        for (int j = 1; j <= 2 && !converged; j++) {
            const double *u = j == 1 ? u1 : u2, *y = j == 1 ? y1 : y2;
            const int m = 2 * it - 2 + j;
            const double f = theta * theta * eta / alpha;
#pragma omp parallel for
            for (int k = 0; k < n; k++) {
                w[k] -= alpha * u[k];
                d[k] = y[k] + f * d[k];
            }
            theta = sqrt(dot(n, w, w)) / tau;
            const double c = 1.0 / sqrt(1.0 + theta * theta);
            tau = tau * theta * c;
            eta = c * c * alpha;
#pragma omp parallel for
            for (int k = 0; k < n; k++) x[k] += eta * d[k];
            if (tau * sqrt((double)m + 1.0) < tol * b_norm) {
                presidue(op, x, b, r);
                resid = sqrt(dot(n, r, r)) / b_norm;
                converged = resid < tol;
            }
        }
        if (converged) break;
        const double rho_new = dot(n, rt, w), beta = rho_new / rho;
        if (!(fabs(rho_new) > 0.0) || !isfinite(beta)) goto restart;
        rho = rho_new;
#pragma omp parallel for
        for (int k = 0; k < n; k++) y1[k] = w[k] + beta * y2[k];
        pmatvec(op, y1, u1);
#pragma omp parallel for
        for (int k = 0; k < n; k++) v[k] = u1[k] + beta * (u2[k] + beta * v[k]);
    }
    presidue(op, x, b, r);
    resid = sqrt(dot(n, r, r)) / b_norm;
    *iters = it;
    *res = resid;
    free(rt); free(w); free(y1); free(y2); free(u1); free(u2); free(v); free(d); free(r);
    return resid < tol;
}

/* BiCGStab(l) (Sleijpen & Fokkema 1993) on M^{-1} A x = M^{-1} b: l BiCG
 * steps, then a minimal-residual polynomial of degree l over the l + 1
 * residuals (modified Gram-Schmidt).  One iteration is one BiCG step (two
 * operator applications, as one BiCGStab iteration); convergence is checked
 * after each cycle of l.  On a breakdown (rho0 = 0, (u_{j+1}, r~) = 0 or
 * sigma_j = 0) it restarts from the current x.  l = 1 is BiCGStab. */
static void axpy(int n, double a, const double *x, double *y)
{
#pragma omp parallel for
    for (int k = 0; k < n; k++) y[k] += a * x[k];
}
static int bicgstabl(krylov_op *op, const double *b, double *x, double tol, int maxit, int ell, int *iters,
                     double *res)
{
    const int n = op->n;
    if (ell < 1) ell = 1;
    if (ell > 8) ell = 8;
    double *R = malloc(sizeof(double) * n * (ell + 1)), *U = malloc(sizeof(double) * n * (ell + 1)),
           *rt = malloc(sizeof(double) * n);
    double tau[9][9], sigma[9], gp[9], g[9], gpp[9];
    prec_apply(op, b, rt);
    double b_norm = sqrt(dot(n, rt, rt));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    double resid, rho0 = 1, alpha = 0, omega = 1;
restart:
    presidue(op, x, b, R);
    memcpy(rt, R, sizeof(double) * n);
    memset(U, 0, sizeof(double) * n);
    rho0 = 1, alpha = 0, omega = 1;
    resid = sqrt(dot(n, R, R)) / b_norm;
    while (resid >= tol && it < maxit && isfinite(resid)) {
        rho0 = -omega * rho0;
        for (int j = 0; j < ell; j++) {
            it++;
            const double rho1 = dot(n, R + (long)j * n, rt), beta = alpha * rho1 / rho0;
            if (!(fabs(rho0) > 0.0) || !isfinite(beta)) goto restart;
            rho0 = rho1;
#pragma omp parallel for
            for (int k = 0; k < n; k++)
                for (int i = 0; i <= j; i++) U[(long)i * n + k] = R[(long)i * n + k] - beta * U[(long)i * n + k];
            pmatvec(op, U + (long)j * n, U + (long)(j + 1) * n);
            const double gam = dot(n, U + (long)(j + 1) * n, rt);
            alpha = rho0 / gam;
            if (!(fabs(gam) > 0.0) || !isfinite(alpha)) goto restart;
#pragma omp parallel for
            for (int k = 0; k < n; k++)
                for (int i = 0; i <= j; i++) R[(long)i * n + k] -= alpha * U[(long)(i + 1) * n + k];
            pmatvec(op, R + (long)j * n, R + (long)(j + 1) * n);
            axpy(n, alpha, U, x);
        }
        for (int j = 1; j <= ell; j++) {
            double *rj = R + (long)j * n;
            for (int i = 1; i < j; i++) {
                tau[i][j] = dot(n, rj, R + (long)i * n) / sigma[i];
                axpy(n, -tau[i][j], R + (long)i * n, rj);
            }
            sigma[j] = dot(n, rj, rj);
            if (!(sigma[j] > 0.0)) goto restart;
            gp[j] = dot(n, R, rj) / sigma[j];
        }
        g[ell] = gp[ell];
        omega = g[ell];
        for (int j = ell - 1; j >= 1; j--) {
            double t = gp[j];
            for (int i = j + 1; i <= ell; i++) t -= tau[j][i] * g[i];
            g[j] = t;
        }
        for (int j = 1; j < ell; j++) {
            double t = g[j + 1];
            for (int i = j + 1; i < ell; i++) t += tau[j][i] * g[i + 1];
            gpp[j] = t;
        }
        axpy(n, g[1], R, x);
        axpy(n, -gp[ell], R + (long)ell * n, R);
        axpy(n, -g[ell], U + (long)ell * n, U);
        for (int j = 1; j < ell; j++) {
            axpy(n, -g[j], U + (long)j * n, U);
            axpy(n, gpp[j], R + (long)j * n, x);
            axpy(n, -gp[j], R + (long)j * n, R);
        }
        resid = sqrt(dot(n, R, R)) / b_norm;
    }
    *iters = it;
    *res = resid;
    free(R); free(U); free(rt);
    return resid < tol;
}

/* the shadow space of IDR(s): s columns of a fixed pseudo-random matrix
 * (an integer hash of (row, column) mapped to [-1, 1) with 24 bits, so the
 * values are exact in any precision), orthonormalised by modified
 * Gram-Schmidt */
double or_idr_shadow_value(long i, int k)
{
    uint32_t h = (uint32_t)i * 0x9E3779B1u + (uint32_t)(k + 1) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
}

/* IDR(s) with biorthogonalisation (van Gijzen & Sonneveld 2011, Alg. 913) on
 * M^{-1} A x = M^{-1} b, with the "maintaining the convergence" choice of
 * omega (angle 0.7).  One iteration is one operator application (s + 1 per
 * cycle); convergence is checked after each.  On a breakdown (M_kk = 0 or
 * t = 0) it restarts from the current x. */
static int idrs(krylov_op *op, const double *b, double *x, double tol, int maxit, int s, int *iters, double *res)
{
    const int n = op->n;
    if (s < 1) s = 1;
    if (s > 8) s = 8;
    double *P = malloc(sizeof(double) * n * s), *G = malloc(sizeof(double) * n * s),
           *U = malloc(sizeof(double) * n * s), *r = malloc(sizeof(double) * n), *v = malloc(sizeof(double) * n),
           *t = malloc(sizeof(double) * n);
    double Mm[8][8], f[8], c[8];
    for (int k = 0; k < s; k++) {
        double *pk = P + (long)k * n;
        for (long i = 0; i < n; i++) pk[i] = or_idr_shadow_value(i, k);
        for (int q = 0; q < k; q++) axpy(n, -dot(n, pk, P + (long)q * n), P + (long)q * n, pk);
        const double nrm = sqrt(dot(n, pk, pk));
        for (long i = 0; i < n; i++) pk[i] /= nrm;
    }
    prec_apply(op, b, r);
    double b_norm = sqrt(dot(n, r, r));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    double resid, om;
restart:
    presidue(op, x, b, r);
    memset(G, 0, sizeof(double) * n * s);
    memset(U, 0, sizeof(double) * n * s);
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) Mm[i][j] = i == j;
    om = 1.0;
    resid = sqrt(dot(n, r, r)) / b_norm;
    while (resid >= tol && it < maxit && isfinite(resid)) {
        for (int i = 0; i < s; i++) f[i] = dot(n, P + (long)i * n, r);
        for (int k = 0; k < s && resid >= tol && it < maxit; k++) {
            /* c = M(k:s, k:s) \ f(k:s), M lower triangular */
            for (int i = k; i < s; i++) {
                double acc = f[i];
                for (int j = k; j < i; j++) acc -= Mm[i][j] * c[j];
                c[i] = acc / Mm[i][i];
            }
            memcpy(v, r, sizeof(double) * n);
            for (int i = k; i < s; i++) axpy(n, -c[i], G + (long)i * n, v);
            double *uk = U + (long)k * n, *gk = G + (long)k * n;
#pragma omp parallel for
            for (long q = 0; q < n; q++) {
                double acc = om * v[q];
                for (int i = k; i < s; i++) acc += c[i] * U[(long)i * n + q];
                uk[q] = acc;
            }
            pmatvec(op, uk, gk);
            it++;
            for (int i = 0; i < k; i++) {
                const double al = dot(n, P + (long)i * n, gk) / Mm[i][i];
                axpy(n, -al, G + (long)i * n, gk);
                axpy(n, -al, U + (long)i * n, uk);
            }
            for (int i = k; i < s; i++) Mm[i][k] = dot(n, P + (long)i * n, gk);
            if (!(fabs(Mm[k][k]) > 0.0)) goto restart;
            const double beta = f[k] / Mm[k][k];
            axpy(n, -beta, gk, r);
            axpy(n, beta, uk, x);
            resid = sqrt(dot(n, r, r)) / b_norm;
            for (int i = k + 1; i < s; i++) f[i] -= beta * Mm[i][k];
        }
        if (resid < tol || it >= maxit || !isfinite(resid)) break;
        memcpy(v, r, sizeof(double) * n);
        pmatvec(op, v, t);
        it++;
        const double tt = dot(n, t, t), tr = dot(n, t, r);
        if (!(tt > 0.0)) goto restart;
        om = tr / tt;
        const double rho = fabs(tr) / (sqrt(tt) * sqrt(dot(n, r, r)));
        if (rho < 0.7) om *= 0.7 / rho;
        axpy(n, -om, t, r);
        axpy(n, om, v, x);
        resid = sqrt(dot(n, r, r)) / b_norm;
    }
    *iters = it;
    *res = resid;
    free(P); free(G); free(U); free(r); free(v); free(t);
    return resid < tol;
}

/* SOR (TNL::Solvers::Linear::SOR): forward sweeps
 *   x_i <- (1 - w) x_i + w (b_i - sum_{j != i} a_ij x_j) / a_ii
 * in row order, each row using the values already updated in this sweep for
 * j < i; w = `omega`.  The preconditioner is not used; the stopping test is
 * the unpreconditioned ||b - A x|| / ||b|| < tol, checked after each sweep
 * (one iteration). */
static int sor(krylov_op *op, const double *b, double *x, double tol, int maxit, double omega, int *iters,
               double *res)
{
    const int n = op->n;
    const int *rp = op->rp, *col = op->col;
    const double *val = op->val;
    double *r = malloc(sizeof(double) * n);
    double b_norm = sqrt(dot(n, b, b));
    if (b_norm == 0.0) b_norm = 1.0;
    or_spmv(n, rp, col, val, x, r);
    for (int k = 0; k < n; k++) r[k] = b[k] - r[k];
    double resid = sqrt(dot(n, r, r)) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && isfinite(resid)) {
        for (int i = 0; i < n; i++) {
            double acc = b[i], d = 0.0;
            for (int q = rp[i]; q < rp[i + 1]; q++) {
                if (col[q] == i) d = val[q];
                else acc -= val[q] * x[col[q]];
            }
            x[i] = (1.0 - omega) * x[i] + omega * acc / d;
        }
        or_spmv(n, rp, col, val, x, r);
        for (int k = 0; k < n; k++) r[k] = b[k] - r[k];
        resid = sqrt(dot(n, r, r)) / b_norm;
        it++;
    }
    *iters = it;
    *res = resid;
    free(r);
    return resid < tol;
}

/* Stationary (Jacobi / Richardson) iteration x += omega M^{-1} (b - A x),
 * omega = 1; with the "diagonal" preconditioner this is the Jacobi method */
static int richardson(krylov_op *op, const double *b, double *x, double tol, int maxit, int *iters,
                      double *res)
{
    const int n = op->n;
    double *r = malloc(sizeof(double) * n);
    prec_apply(op, b, r);
    double b_norm = sqrt(dot(n, r, r));
    if (b_norm == 0.0) b_norm = 1.0;
    presidue(op, x, b, r);
    double resid = sqrt(dot(n, r, r)) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && isfinite(resid)) {
#pragma omp parallel for
        for (int k = 0; k < n; k++) x[k] += r[k];
        presidue(op, x, b, r);
        resid = sqrt(dot(n, r, r)) / b_norm;
        it++;
    }
    *iters = it;
    *res = resid;
    free(r);
    return resid < tol;
}

int or_krylov(int nrows, const int *rp, const int *col, const double *val, const double *b,
              double *x, int solver, int precond, double tol, int max_iter, int gmres_restart,
              int *iterations, double *residue)
{
    krylov_op op = {nrows, rp, col, val, precond, NULL, NULL, NULL, NULL, NULL};
    op.tmp = malloc(sizeof(double) * nrows);
    int rc = 0;
    if (precond == OR_PREC_DIAGONAL) {
        /* Diagonal preconditioner: inverse of the matrix diagonal */
        op.dinv = malloc(sizeof(double) * nrows);
        for (int r = 0; r < nrows; r++) {
            double dg = 0;
            for (int k = rp[r]; k < rp[r + 1]; k++)
                if (col[k] == r) dg = val[k];
            op.dinv[r] = dg != 0.0 ? 1.0 / dg : 1.0;
        }
    } else if (precond == OR_PREC_ILU0) {
        const long nnz = rp[nrows];
        op.lu = malloc(sizeof(double) * nnz);
        memcpy(op.lu, val, sizeof(double) * nnz);
        if (or_ilu0_factorize(nrows, rp, col, op.lu)) rc = -1;
    } else if (precond == OR_PREC_ILUT) {
        if (or_ilut_factorize(nrows, rp, col, val, ilut_p, ilut_tau, &op.frp, &op.fcol, &op.lu) < 0) rc = -1;
    }
    int conv = 0;
    if (!rc) {
        if (solver == OR_GMRES) conv = gmres(&op, b, x, tol, max_iter, gmres_restart, iterations, residue);
        else if (solver == OR_CG) conv = cg(&op, b, x, tol, max_iter, iterations, residue);
        else if (solver == OR_TFQMR) conv = tfqmr(&op, b, x, tol, max_iter, iterations, residue);
        else if (solver == OR_JACOBI) conv = richardson(&op, b, x, tol, max_iter, iterations, residue);
        else if (solver == OR_BICGSTABL) conv = bicgstabl(&op, b, x, tol, max_iter, gmres_restart, iterations, residue);
        else if (solver == OR_IDRS) conv = idrs(&op, b, x, tol, max_iter, gmres_restart, iterations, residue);
        else if (solver == OR_SOR) conv = sor(&op, b, x, tol, max_iter, gmres_restart / 100.0, iterations, residue);
        else conv = bicgstab(&op, b, x, tol, max_iter, iterations, residue);
    }
    free(op.tmp); free(op.dinv); free(op.lu); free(op.frp); free(op.fcol);
    return rc ? -1 : conv;
}

/* Solver::solveLinearSystem (Solver.hpp:1230-1265): initial guess = current Z_iF */
int or_solve(or_problem *p, int solver, int precond, double tol, int max_iter, int gmres_restart,
             int *iterations, double *residue)
{
    int rc = or_krylov(p->nrows, p->rowptr, p->colidx, p->val, p->rhs, p->Z_iF, solver, precond, tol,
                       max_iter, gmres_restart, iterations, residue);
    if (rc != 1)
        snprintf(p->err, sizeof p->err,
                 "MHFEM error: the linear system solver did not converge (%d iterations).", *iterations);
    return rc;
}

int or_rows(const or_problem *p) { return p->nrows; }
long or_nnz(const or_problem *p) { return p->nnz; }
int *or_rowptr(or_problem *p) { return p->rowptr; }
int *or_colidx(or_problem *p) { return p->colidx; }
double *or_values(or_problem *p) { return p->val; }
