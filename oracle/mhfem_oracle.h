/*
 * mhfem_oracle.h -- CPU oracle for the MHFEM per-time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the host path of
 * TNL-MHFEM (the headers under include/TNL/MHFEM) used as the checker for the HIP
 * implementation.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never does.
 *
 * Array layouts are the reference's CUDA permutations (BaseModel.h:50-124):
 *   Z_iF        [F][n]            (E*n + i)
 *   Z_iK, m_iK, f_iK, R_iK        [n][C]
 *   N_ijK, D_ijK, r_ijK           [n][n][C]
 *   u_ijKe, a_ijKe, R_ijKe        [n][n][FPC][C]
 *   w_iKe, v_iKe                  [n][FPC][C]
 *   m_iE_upw    [F][n]
 *   Z_ijE_upw   [F][n(j)][n(i)]
 *   b_ijK_storage                 [n][n][S][C]   (S = 9 triangles, 14 tetrahedra,
 *                                                  1 edges, 2 rectangles, 3 voxels)
 * Boundary-condition arrays use the (i,E) layout i*F + E
 * (BoundaryConditions.hpp:295, Stationary.h:22,36).
 *
 * Parity status of each part is stated at its definition in mhfem_oracle.c.
 */
#ifndef MHFEM_ORACLE_H
#define MHFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_problem or_problem;

/* advection discretisations (BaseModel.h:12-16) */
enum { OR_EXPLICIT_UPWIND = 0, OR_IMPLICIT_UPWIND = 1, OR_IMPLICIT_TRACE = 2 };
/* boundary condition tags (BoundaryConditionsType.h:7-31) */
enum { OR_FIXED_VALUE = 0, OR_FIXED_FLUX = 1, OR_ADVECTIVE_OUTFLOW = 2, OR_FIXED_FLUX_NO_ADVECTION = 3 };
/* linear solvers / preconditioners */
enum { OR_BICGSTAB = 0, OR_GMRES = 1, OR_CG = 2, OR_TFQMR = 3, OR_JACOBI = 4, OR_BICGSTABL = 5, OR_IDRS = 6,
       OR_SOR = 7 };
/* or_krylov's gmres_restart is the restart length of GMRES, l of BiCGStab(l),
 * s of IDR(s) and 100 w for SOR's relaxation w; or_idr_shadow_value: entry
 * (i, k) of IDR(s)'s shadow space before orthonormalisation */
double or_idr_shadow_value(long i, int k);
enum { OR_PREC_NONE = 0, OR_PREC_DIAGONAL = 1, OR_PREC_ILU0 = 2, OR_PREC_ILUT = 3 };
/* ILUT(p, tau) factors (CSR, sorted columns, unit-lower L left of U's
 * diagonal; arrays malloc'ed) and the p, tau that or_krylov's ILUT uses */
long or_ilut_factorize(int n, const int *rp, const int *col, const double *val, int p, double tau, int **rp_out,
                       int **col_out, double **val_out);
void or_set_ilut(int p, double tau);

/* Topology only: face numbering and incidence.  Returns the number of faces
 * (>= 0) or -1 on a non-conforming mesh.  cell_faces: nc*(dim+1),
 * face_cells: (nc*(dim+1))*2 capacity, second cell = -1 on boundary faces. */
int or_topology(int dim, int nv, const double *points, int nc, const int *cell_vertices,
                int *cell_faces, int *face_cells);

or_problem *or_create(int dim, int nv, const double *points, int nc, const int *cell_vertices,
                      int n, int lumping, int advection, int do_mobility_upwind);
/* cell shapes: simplices (dim+1 vertices) or boxes (2^dim vertices in VTK
 * order: edges, axis-aligned rectangles and voxels); dim 1 is always a box */
enum { OR_SIMPLEX = 0, OR_BOX = 1 };
int or_topology_shape(int dim, int shape, int nv, const double *points, int nc,
                      const int *cell_vertices, int *cell_faces, int *face_cells);
or_problem *or_create_shape(int dim, int shape, int nv, const double *points, int nc,
                            const int *cell_vertices, int n, int lumping, int advection,
                            int do_mobility_upwind);
void or_destroy(or_problem *p);
/* OpenMP threads of the oracle's parallel loops */
void or_set_threads(int threads);
const char *or_error(const or_problem *p);

int or_num_cells(const or_problem *p);
int or_num_faces(const or_problem *p);
/* pointer to an internal array (names as in the header comment, plus
 * "cell_measure", "face_measure", "tags"(uint8), "bc_values", "bc_dirichlet",
 * "bc_mobility", "cell_faces"(int), "face_cells"(int), "rhs"); size in elements */
void *or_array(or_problem *p, const char *name, long *size);

/* Solver::setInitialCondition tail (Solver.hpp:319-341): Z_iF := average of
 * neighbouring Z_iK, v_iKe := 0 */
void or_init(or_problem *p);
/* Solver::setupLinearSystem (Solver.hpp:448-484): CSR pattern */
int or_setup_linear_system(or_problem *p);
void or_pre_iterate(or_problem *p, double time, double tau);
void or_assemble(or_problem *p, double time, double tau);
/* returns 1 if converged; iterations/residue through the pointers */
int or_solve(or_problem *p, int solver, int precond, double tol, int max_iter, int gmres_restart,
             int *iterations, double *residue);
void or_post_iterate(or_problem *p, double time, double tau);

/* matrix access (after or_setup_linear_system) */
int or_rows(const or_problem *p);
long or_nnz(const or_problem *p);
int *or_rowptr(or_problem *p);
int *or_colidx(or_problem *p);
double *or_values(or_problem *p);

/* standalone arithmetic (used by tests to pin kernels in isolation) */
/* ILU(0) of a CSR matrix with sorted columns, in place on val (Saad Alg. 10.4);
 * returns 0 or -1 on a zero pivot / missing diagonal */
int or_ilu0_factorize(int nrows, const int *rowptr, const int *col, double *val);
void or_ilu0_apply(int nrows, const int *rowptr, const int *col, const double *lu,
                   const double *x, double *y);
void or_spmv(int nrows, const int *rowptr, const int *col, const double *val, const double *x,
             double *y);
/* Krylov on an explicit CSR system (same code path as or_solve) */
int or_krylov(int nrows, const int *rowptr, const int *col, const double *val, const double *b,
              double *x, int solver, int precond, double tol, int max_iter, int gmres_restart,
              int *iterations, double *residue);
/* geometric inverse mass matrix of one cell, packed + row sums (MassMatrix.h) */
void or_local_mass_inverse(int dim, int lumping, const double *cell_points, double measure,
                           double D, double *out);

#ifdef __cplusplus
}
#endif
#endif
