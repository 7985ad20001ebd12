/*
 * mhfem.h -- C ABI of the MI355X-native MHFEM per-time-step path.
 *
 * One context (mhfem_ctx) owns one mesh, one model (n equations) and every
 * device buffer of the scheme, the linear system and the Krylov solver, on
 * the current HIP device, and launches all its kernels on one HIP stream.
 * The entry points replace the per-step methods of the reference's
 * TNL::MHFEM::Solver (include/TNL/MHFEM/Solver.h:64-117) instantiated with
 * Device = TNL::Devices::Cuda; the C++ facade in include/mhfem/ mirrors that
 * class and control.h on top of them.
 *
 * Conventions.  All pointers passed in are HOST pointers unless the name says
 * "device".  Arrays use the reference's CUDA permutations (BaseModel.h:50-124):
 *   Z_iF, m_iE_upw : E*n + i          Z_iK, m_iK, f_iK, R_iK : i*C + K
 *   N_ijK, D_ijK, r_ijK : (i*n+j)*C + K
 *   u_ijKe, a_ijKe, R_ijKe : ((i*n+j)*FPC + e)*C + K
 *   w_iKe, v_iKe : (i*FPC + e)*C + K   Z_ijE_upw : (E*n + j)*n + i
 *   b_ijK_storage : ((i*n+j)*S + s)*C + K   (S = 9 triangles, 14 tetrahedra)
 * Three derived arrays are not kept in these layouts between calls, and
 * mhfem_get_array / mhfem_device_array write them when asked (a snapshot):
 * b_ijK_storage (= M^{-1} D_ij, MassMatrix.h:586-597) from the per-cell
 * M^{-1} of the fixed geometry and the current D_ijK; R_ijKe and R_iK from
 * the per-cell records the last mhfem_pre_iterate wrote, which keep them in
 * the order the face assembly reads them.
 * Boundary-condition arrays are indexed i*F + E (BoundaryConditions.hpp:295).
 * Every function returns 0 on success or a negative MHFEM_E_* code; the
 * message is available from mhfem_last_error().  A context is not
 * thread-safe; distinct contexts are independent.
 */
#ifndef MHFEM_H
#define MHFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhfem_ctx mhfem_ctx;

enum {
    MHFEM_OK = 0,
    MHFEM_E_ARGUMENT = -1,
    MHFEM_E_HIP = -2,
    MHFEM_E_MESH = -3,
    MHFEM_E_STATE = -4,
    MHFEM_E_NOT_CONVERGED = -5,
    MHFEM_E_BREAKDOWN = -6,
    MHFEM_E_FILE = -7
};

/* MassLumping (MassMatrix.h:16-19), AdvectionDiscretization (BaseModel.h:12-16) */
enum { MHFEM_LUMPING_DISABLED = 0, MHFEM_LUMPING_ENABLED = 1 };
enum { MHFEM_EXPLICIT_UPWIND = 0, MHFEM_IMPLICIT_UPWIND = 1, MHFEM_IMPLICIT_TRACE = 2 };
/* BoundaryConditionsType (BoundaryConditionsType.h:7-31) */
enum {
    MHFEM_BC_FIXED_VALUE = 0,
    MHFEM_BC_FIXED_FLUX = 1,
    MHFEM_BC_ADVECTIVE_OUTFLOW = 2,
    MHFEM_BC_FIXED_FLUX_NO_ADVECTION = 3
};

/* Array identifiers for mhfem_set_array / mhfem_get_array / mhfem_device_array. */
enum {
    MHFEM_Z_iF = 0, MHFEM_Z_iK, MHFEM_N_ijK, MHFEM_u_ijKe, MHFEM_m_iK, MHFEM_D_ijK,
    MHFEM_w_iKe, MHFEM_a_ijKe, MHFEM_r_ijK, MHFEM_f_iK,
    MHFEM_v_iKe, MHFEM_m_iE_upw, MHFEM_Z_ijE_upw, MHFEM_b_ijK_storage, MHFEM_R_ijKe, MHFEM_R_iK,
    MHFEM_BC_VALUES, MHFEM_BC_DIRICHLET, MHFEM_BC_MOBILITY, MHFEM_RHS,
    MHFEM_CELL_MEASURE, MHFEM_FACE_MEASURE,
    MHFEM_NUM_ARRAYS
};

typedef struct {
    int dim;                    /* 2: triangles, 3: tetrahedra */
    int num_points;
    int num_cells;
    const double *points;       /* num_points*dim */
    const int32_t *cell_vertices; /* num_cells*(dim+1); local face e is opposite local vertex e */
} mhfem_mesh_desc;

typedef struct {
    int num_equations;          /* n, 1..3 */
    int mass_lumping;           /* MHFEM_LUMPING_* */
    int advection;              /* MHFEM_*_UPWIND / MHFEM_IMPLICIT_TRACE */
    int do_mobility_upwind;     /* BaseModel::do_mobility_upwind (BaseModel.h:163) */
} mhfem_model_desc;

/* Linear solver selection (configSetup.h:31-62, the names of
 * TNL::Solvers::getLinearSolverOptions).  BiCGStab and GMRES(m) as in the
 * reference; CG for symmetric positive definite systems; TFQMR, restarted
 * from the current iterate on a breakdown; JACOBI the stationary iteration
 * x += M^{-1}(b - A x), which is the Jacobi method on the unit-diagonal MHFEM
 * rows; BICGSTABL BiCGStab(l) (Sleijpen & Fokkema), one iteration per BiCG
 * step, convergence checked per cycle of l; IDRS IDR(s) with
 * biorthogonalisation (van Gijzen & Sonneveld, Alg. 913), one iteration per
 * operator application, on a fixed pseudo-random orthonormal shadow space.
 * All are left preconditioned and stop on
 * ||M^{-1}(b - A x)|| / ||M^{-1} b|| < convergence_residue.  SOR: forward
 * sweeps with relaxation sor_omega (one iteration per sweep), no
 * preconditioner, stopping on ||b - A x|| / ||b||.
 * The FP32 precisions, batches and graph steps take BiCGStab only. */
enum {
    MHFEM_SOLVER_BICGSTAB = 0,
    MHFEM_SOLVER_GMRES = 1,
    MHFEM_SOLVER_CG = 2,
    MHFEM_SOLVER_TFQMR = 3,
    MHFEM_SOLVER_JACOBI = 4,
    MHFEM_SOLVER_BICGSTABL = 5,
    MHFEM_SOLVER_IDRS = 6,
    MHFEM_SOLVER_SOR = 7
};
/* MHFEM_PREC_AMG: algebraic multigrid in the configuration of the
 * reference's Ginkgo branch (Solver.hpp:1025-1075): parallel graph matching
 * aggregation, one V-cycle with 2 (coarsest: 4) Richardson sweeps of relaxation
 * 0.9 preconditioned by block Jacobi (blocks of at most n rows), rebuilt when
 * the last solve took more than min(2u, u + 5) iterations, u those right after
 * the last rebuild (Solver.hpp:1017, 1158-1160).  FP64 solvers, single
 * systems, the four per-step calls (not mhfem_step). */
/* MHFEM_PREC_ILU_ISAI: the Ginkgo branch's "ILU_ISAI" (Solver.hpp:1089-1130):
 * ILU(0) factors replaced by incomplete sparse approximate inverses on the
 * patterns of L^2 and U^2, applied as two sparse products; every row of the
 * inverses is solved exactly.  MHFEM_PREC_PARILU_ISAI ("PARILU_ISAI"): the
 * same with ParILU factors (5 synchronous fixed-point sweeps of the ILU(0)
 * equations instead of the level-scheduled elimination).  Same reuse rule,
 * configurations and limits as MHFEM_PREC_AMG.  (PARILUT_ISAI is not
 * provided.) */
/* MHFEM_PREC_ILUT: TNL's "ilut" (configSetup.h:61), ILUT(p, tau) after Saad
 * (Alg. 10.6), factorised on the host every step and applied by level-
 * scheduled triangular solves on its own pattern; FP64 solvers, single
 * systems, the four per-step calls, not with SOR. */
enum {
    MHFEM_PREC_NONE = 0,
    MHFEM_PREC_DIAGONAL = 1,
    MHFEM_PREC_ILU0 = 2,
    MHFEM_PREC_AMG = 3,
    MHFEM_PREC_ILU_ISAI = 4,
    MHFEM_PREC_PARILU_ISAI = 5,
    MHFEM_PREC_ILUT = 6
};
/* precision of the Krylov iteration.  FP32_REFINED: FP64 iterative
   refinement whose every correction is an FP32 BiCGStab solve started from
   zero to `inner_residue`.  FP32_RELIABLE: one FP32 BiCGStab recurrence whose
   residual is replaced by the FP64 residual (and the FP32 correction added
   to the FP64 solution) each time it has dropped by `inner_residue`; the
   search direction is kept across the replacements. */
enum { MHFEM_PRECISION_FP64 = 0, MHFEM_PRECISION_FP32_REFINED = 1, MHFEM_PRECISION_FP32_RELIABLE = 2 };


typedef struct {
    int solver;                 /* MHFEM_SOLVER_* ("linear-solver") */
    int preconditioner;         /* MHFEM_PREC_*   ("preconditioner") */
    double convergence_residue; /* "convergence-residue" */
    int max_iterations;         /* "max-iterations" */
    int gmres_restart;          /* "gmres-restarting-max" */
    int precision;              /* MHFEM_PRECISION_* */
    double inner_residue;       /* FP32 inner tolerance (FP32_REFINED) or residual drop per
                                   FP64 update (FP32_RELIABLE) */
    int check_every;            /* host convergence poll period (iterations), >= 1 */
    int rows_per_block;         /* rows per work-group of the iteration kernels, 0: default
                                   (mhfem_autotune picks it by measurement) */
    int face_groups;            /* BiCGStab iterations with Jacobi or no preconditioner on a
                                   single system (FP32 and FP64): 1 keeps a copy of the matrix
                                   and the iteration's vectors in an order that groups faces of
                                   one type inside windows of 4096 faces, for coalesced gathers
                                   (results unchanged up to rounding); 0 (default): natural order */
    int bicgstab_ell;           /* l of BiCGStab(l) ("bicgstab-ell"), 1..8; 0: 1 */
    int idr_s;                  /* s of IDR(s) ("idr-s"), 1..8; 0: 4 */
    double sor_omega;           /* relaxation of SOR ("sor-omega"), in (0, 2); 0: 1 */
    int ilut_p;                 /* ILUT: largest entries kept per row in L and in U ("ilut-p"); 0: 10 */
    double ilut_threshold;      /* ILUT: drop tolerance relative to the row norm ("ilut-threshold");
                                   0: 1e-4 */
} mhfem_solver_config;

typedef struct {
    int iterations;             /* Krylov iterations of the last solve (inner, for IR) */
    int outer_iterations;       /* refinement steps (1 for plain FP64) */
    double residue;             /* final relative (preconditioned) residue */
    int converged;
    int fallback_iterations;    /* FP64 BiCGStab iterations after the FP32 iteration stopped
                                   reducing the FP64 residual (mixed precision only) */
} mhfem_solve_stats;

/* Per-phase device times (ms, accumulated) under the reference's timer names
 * (Solver.hpp:1353-1386). */
typedef struct {
    double pre_iterate, update_b_R_Q, upwind, assemble, preconditioner, linear_solver,
        post_iterate;
    long long iterations, time_steps;
    /* timing level 2: every 9th SpMV launch of the Krylov iteration is
       bracketed by HIP events on the context's stream */
    double spmv_ms;       /* summed duration of the timed SpMV launches that did the
                             product (launches queued after convergence return at once
                             and are left out: below a quarter of the median) */
    long long spmv_timed; /* number of those launches */
} mhfem_timers;

/* ---- lifetime (Solver::setMesh, Solver.hpp:26-81; BaseModel::allocate, BaseModel.h:168-195) */
mhfem_ctx *mhfem_create(const mhfem_mesh_desc *mesh, const mhfem_model_desc *model);
/* Every cell shape of the reference's examples/config.h:16-20:
 * vertices_per_cell 2 (edges, dim 1), 3 (triangles), 4 with dim 2
 * (quadrangles: axis-aligned rectangles, vertices counter-clockwise from the
 * lower-left one as VTK orders them, MassMatrix.h:146-303), 4 with dim 3
 * (tetrahedra) or 8 (hexahedra: axis-aligned voxels in VTK order,
 * MassMatrix.h:305-508).  cell_vertices then holds
 * num_cells*vertices_per_cell entries; b_ijK_storage has MassMatrix::size
 * values per (i,j,K): 9 or 14 for simplices, 1, 2 or 3 for edges, rectangles
 * and voxels.  mhfem_create(mesh, model) is
 * mhfem_create_cells(mesh, mesh->dim + 1, model). */
mhfem_ctx *mhfem_create_cells(const mhfem_mesh_desc *mesh, int vertices_per_cell, const mhfem_model_desc *model);
void mhfem_destroy(mhfem_ctx *ctx);
const char *mhfem_last_error(const mhfem_ctx *ctx); /* ctx may be NULL: error of the last create */

/* ---- topology (mesh_helpers.h:49-265) */
int mhfem_get_counts(const mhfem_ctx *ctx, int *cells, int *faces, int *boundary_faces);
/* host copies: cell_faces[C*FPC], face_cells[F*2] (second = -1 on the boundary) */
int mhfem_get_topology(const mhfem_ctx *ctx, int32_t *cell_faces, int32_t *face_cells);

/* ---- model data (BaseModel.h:266-286, UniformCoefficient.h:7-53):
 * count == full size uploads the array, count == 1 makes a coefficient
 * uniform (one broadcast value). */
int mhfem_set_array(mhfem_ctx *ctx, int which, const double *host, long count);
int mhfem_get_array(mhfem_ctx *ctx, int which, double *host, long count);
long mhfem_array_size(const mhfem_ctx *ctx, int which);
/* device pointer of an array (for in-place updates by model hooks on the
 * context's stream); NULL for uniform coefficients */
void *mhfem_device_array(mhfem_ctx *ctx, int which);
void *mhfem_stream(mhfem_ctx *ctx); /* the hipStream_t of the context */

/* ---- boundary conditions (BoundaryConditions::init, BoundaryConditions.hpp:272-281;
 * BoundaryConditionsStorage.h:9-15); arrays of n*F in the i*F+E layout */
int mhfem_set_boundary_conditions(mhfem_ctx *ctx, const uint8_t *tags, const double *values,
                                  const double *dirichlet_values);

/* ---- boundary model: the BoundaryModel argument of the reference's Solver
 * (Solver.h:28-31), whose getDirichletValue / getNeumannValue (i, E, time,
 * tau) the assembly calls on the device with time = t + tau (Solver.hpp:948,
 * BoundaryConditions.hpp:379, 414, 455).  The default is the reference's
 * Stationary model (BoundaryModels/Stationary.h:14-37): the stored
 * dirichlet_values[i*F+E] and values[i*F+E].  A time function per species
 * and per kind of value makes the model time dependent:
 *   getDirichletValue(i, E, time) = dirichlet_values[i*F+E] * f_D,i(time)
 *   getNeumannValue(i, E, time)   = values[i*F+E] * f_N,i(time)
 * evaluated on the device inside the assembly, for the boundary rows only
 * (no per-step upload).  Kinds: CONSTANT f = offset; SINE f = offset +
 * amplitude * sin(angular_frequency * t + phase); TABLE f piecewise linear
 * through (times[k], values[k]), k < points (times strictly ascending),
 * constant before the first and after the last point (the arrays are
 * copied).  which = MHFEM_BC_DIRICHLET or MHFEM_BC_VALUES; f = NULL restores
 * the stationary value of that species. */
enum { MHFEM_TIME_CONSTANT = 0, MHFEM_TIME_SINE = 1, MHFEM_TIME_TABLE = 2 };
typedef struct {
    int kind;                   /* MHFEM_TIME_* */
    double offset, amplitude, angular_frequency, phase;
    int points;                 /* MHFEM_TIME_TABLE */
    const double *times, *values;
} mhfem_time_function;
int mhfem_set_boundary_time_function(mhfem_ctx *ctx, int which, int species, const mhfem_time_function *f);

/* ---- Solver::setInitialCondition (Solver.hpp:255-344) for component i
 * (BaseModel::setInitialCondition, BaseModel.h:316-326) */
int mhfem_set_initial_condition(mhfem_ctx *ctx, int i, const double *Z_K, long count);
/* tail of setInitialCondition: Z_iF := face averages of Z_iK, v_iKe := 0
 * (Solver.hpp:319-341) */
int mhfem_init(mhfem_ctx *ctx);

/* ---- Solver::setup + setupLinearSystem (Solver.hpp:83-253, 346-485):
 * solver/preconditioner choice and the fixed CSR pattern */
int mhfem_setup_linear_system(mhfem_ctx *ctx, const mhfem_solver_config *cfg);

/* ---- one time step (control.h:74-101) */
int mhfem_pre_iterate(mhfem_ctx *ctx, double time, double tau);     /* Solver.hpp:600-892 */
int mhfem_assemble(mhfem_ctx *ctx, double time, double tau);        /* Solver.hpp:894-962 */
int mhfem_solve(mhfem_ctx *ctx, mhfem_solve_stats *stats);          /* Solver.hpp:979-1266 */
int mhfem_post_iterate(mhfem_ctx *ctx, double time, double tau);    /* Solver.hpp:1268-1324 */
/* control.h::solve(problem, t0, t1, dt) (control.h:64-102); stats of the last step */
int mhfem_solve_interval(mhfem_ctx *ctx, double t0, double t1, double dt,
                         mhfem_solve_stats *last);

/* One time step (preIterate, assemble, solve, postIterate) replayed from HIP
 * graphs captured on the first call (re-captured when tau changes): FP64
 * BiCGStab only; equivalent to the four calls above. */
int mhfem_step(mhfem_ctx *ctx, double time, double tau, mhfem_solve_stats *stats);
/* Measure the rows per work-group of the iteration kernels on the assembled
 * pattern and keep the fastest; returns the choice and its kernel time. */
int mhfem_autotune(mhfem_ctx *ctx, int *rows_per_block, double *kernel_ms);

/* ---- batched ensembles (no counterpart in the reference, where each
 * instance is its own Solver).  The meshes are concatenated into one context
 * whose system is block diagonal; every instance (segment) keeps its own
 * BiCGStab scalars and converges independently, and a step of the whole
 * ensemble is one graph launch (mhfem_step).  Arrays are the concatenation of
 * the instances' arrays in the layouts above (cells and faces numbered
 * instance by instance). */
mhfem_ctx *mhfem_create_batch(const mhfem_mesh_desc *meshes, int count, const mhfem_model_desc *model);
/* cell_offsets[count+1], face_offsets[count+1] (either may be NULL) */
int mhfem_batch_info(const mhfem_ctx *ctx, int *count, int32_t *cell_offsets, int32_t *face_offsets);
/* iterations and final residue of every instance in the last solve */
int mhfem_batch_stats(mhfem_ctx *ctx, int32_t *iterations, double *residues, int count);

/* ---- linear system export (Solver::saveLinearSystem, Solver.hpp:964-977):
 * rowptr[rows+1], col[nnz], val[nnz], rhs[rows]; any pointer may be NULL */
int mhfem_system_size(const mhfem_ctx *ctx, int *rows, long *nnz);
int mhfem_export_system(mhfem_ctx *ctx, int32_t *rowptr, int32_t *col, double *val, double *rhs);
/* solve an explicit right-hand side with the current matrix (testing aid) */
int mhfem_solve_rhs(mhfem_ctx *ctx, const double *rhs, double *x, mhfem_solve_stats *stats);
/* mhfem_solve_rhs allocating nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  mhfem_solve_rhs
 * stages rhs and x in device buffers it allocates on every call; this call
 * stages them in a fixed device array of the library instead (2 *
 * MHFEM_NOALLOC_MAX_ROWS doubles, 16 MiB, one per GPU the library runs on,
 * held until the process ends).  Limits: at most MHFEM_NOALLOC_MAX_ROWS rows,
 * and only configurations whose solve allocates nothing: BiCGStab with no,
 * diagonal or ILU(0) preconditioning (any precision), learned guess off;
 * anything else is refused with MHFEM_E_ARGUMENT.  Every call on one GPU
 * uses the same array, which keeps the last call's contents: calls that
 * overlap on one GPU (several host threads or contexts) race, and the caller
 * must serialise them. */
#define MHFEM_NOALLOC_MAX_ROWS 1048576
int mhfem_solve_rhs_noalloc(mhfem_ctx *ctx, const double *rhs, double *x, mhfem_solve_stats *stats);

/* ---- learned initial guess of each solve.  No counterpart in the reference:
 * the reference starts every Krylov solve from the previous step's Z_iF
 * (Solver.hpp:1247-1255); this replaces only that starting vector.  A per-DOF
 * MLP (16 features -> 64 -> 64 -> 1, BF16 MFMA, FP32 accumulation) predicts
 * the increment of Z_iF from 2-hop features of the warm-start residual and of
 * the previous increments; the prediction is used only when it lowers the
 * initial residual, and the Krylov solve still decides convergence. */
/* how the starting vector is formed: the least-residual combination of the
 * previous Z_iF plus the network's predicted increment and the last
 * history_depth increments (COMBINED), of the network's prediction alone
 * (NETWORK), or of the previous increments alone (HISTORY: the method without
 * the network) */
enum { MHFEM_GUESS_COMBINED = 0, MHFEM_GUESS_NETWORK = 1, MHFEM_GUESS_HISTORY = 2 };

typedef struct {
    int enabled;          /* 0: previous Z_iF (reference behaviour), 1: learned guess */
    int train_steps;      /* train online on the first train_steps solves (0: frozen) */
    int epochs;           /* Adam steps per training solve */
    float learning_rate;  /* Adam step size */
    unsigned seed;        /* weight initialisation */
    int combine;          /* MHFEM_GUESS_* */
    int sample_stride;    /* an Adam step uses every sample_stride-th tile of 16 DOFs (>= 1) */
    int reject_after;     /* once training is over, the network is switched off after this many
                             consecutive solves in which its direction lowers the squared
                             initial residual by less than 1 % beyond the previous increments
                             (0: never) */
    int history_depth;    /* previous increments in the combination, 1..16 (0: 2) */
} mhfem_guess_config;

typedef struct {
    int used;             /* the prediction was accepted as the starting vector */
    double r0_norm;       /* ||b - A Z_iF(previous)|| */
    double r_guess_norm;  /* ||b - A x_guess|| (== r0_norm when not evaluated) */
    double loss;          /* mean squared training loss of the last epoch (normalised units) */
    int trained_steps;    /* solves used for training so far */
    double coef_network;  /* weight of the network's direction in the combination */
    double network_gain;  /* relative drop of the squared initial residual due to the network */
    int network_active;   /* the network is evaluated (not switched off by reject_after) */
} mhfem_guess_stats;

int mhfem_guess_configure(mhfem_ctx *ctx, const mhfem_guess_config *cfg);
int mhfem_guess_stats_last(const mhfem_ctx *ctx, mhfem_guess_stats *out);
int mhfem_guess_num_params(void);
/* FP32 master weights (W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3) */
int mhfem_guess_get_weights(mhfem_ctx *ctx, float *params, int count);
int mhfem_guess_set_weights(mhfem_ctx *ctx, const float *params, int count);
/* the network on explicit inputs (features as BF16 bit patterns [rows][16]):
 * forward y[rows], and the gradient of mean((y - t)^2) (count+1 floats: the
 * last one is the summed squared error) */
int mhfem_guess_forward(mhfem_ctx *ctx, const uint16_t *features, long rows, float *y);
int mhfem_guess_gradient(mhfem_ctx *ctx, const uint16_t *features, const float *targets, long rows,
                         float *grad);

/* ---- files on either side of the path (host code; no GPU work).  Errors
 * leave their message for mhfem_last_error(NULL). */
/* Mesh files of lines, triangles, quadrangles, tetrahedra or hexahedra (one
 * shape per file) with their cell data: VTK XML (.vtu:
 * ascii, base64 binary with or without zlib compression, raw or base64
 * appended data) and legacy VTK (.vtk: ASCII or BINARY).  Replace
 * TNL::Meshes::loadMesh (control.h:20-48) and the reader's readCellData for
 * "InitialCondition[Z<i>]" (Solver.hpp:296-306).  format: "auto" (by the
 * extension, as control.h:24-31 does), "vtu" or "vtk".  NULL on error. */
typedef struct mhfem_mesh_file mhfem_mesh_file;
mhfem_mesh_file *mhfem_mesh_file_open(const char *path, const char *format);
void mhfem_mesh_file_close(mhfem_mesh_file *f);
/* dimension, vertices per cell (the arguments of mhfem_create_cells), counts
 * of points and cells */
int mhfem_mesh_file_info(const mhfem_mesh_file *f, int *dim, int *vertices_per_cell, long *points, long *cells);
/* points[points*dim], cell_vertices[cells*vertices_per_cell] in the file's vertex order */
int mhfem_mesh_file_read(const mhfem_mesh_file *f, double *points, int32_t *cell_vertices);
/* a cell-data array by name (converted to double, components interleaved);
 * returns its length (out may be NULL to ask for it) or a negative code */
long mhfem_mesh_file_cell_data(const mhfem_mesh_file *f, const char *name, double *out, long capacity);
int mhfem_mesh_file_cell_data_count(const mhfem_mesh_file *f);
const char *mhfem_mesh_file_cell_data_name(const mhfem_mesh_file *f, int index);

/* VTU writer of snapshots (HeatEquationModel.h:164-231: TNL's VTUWriter,
 * FileFormat::zlib_compressed): FieldData CYCLE (Int32) and TIME (Float64)
 * when >= 0, the mesh, and `fields` cell arrays of num_cells doubles. */
enum { MHFEM_VTK_ASCII = 0, MHFEM_VTK_BINARY = 1, MHFEM_VTK_ZLIB = 2 };
int mhfem_write_vtu(const char *path, const mhfem_mesh_desc *mesh, int vertices_per_cell, long cycle, double time,
                    int fields, const char *const *names, const double *const *values, int format);

/* BoundaryConditionsStorage in TNL's binary object format
 * (BoundaryConditionsStorage.h:22-72): type "TNL::MHFEM::BoundaryConditionsStorage< double >",
 * dofSize, then the tags, values and dirichletValues arrays (i*F + E).  Load
 * with tags, values and dirichlet NULL to learn dof_size first. */
int mhfem_bc_save(const char *path, int64_t dof_size, const uint8_t *tags, const double *values,
                  const double *dirichlet);
int mhfem_bc_load(const char *path, int64_t *dof_size, uint8_t *tags, double *values, double *dirichlet,
                  int64_t capacity);
/* a vector as TNL::Containers::Array< double > (Solver.hpp:974-975); load
 * returns the length (data may be NULL) or a negative code */
int mhfem_save_vector(const char *path, const double *data, int64_t count);
long mhfem_load_vector(const char *path, double *data, int64_t capacity);
/* Solver::saveLinearSystem (Solver.hpp:966-977), used when a solve fails:
 * <directory>/matrix.mtx (Matrix Market coordinate, 1-based), dof.vec.tnl
 * (the current Z_iF) and rhs.vec.tnl */
int mhfem_save_linear_system(mhfem_ctx *ctx, const char *directory);

/* ---- instrumentation (Solver::writeEpilog, Solver.hpp:1326-1391) */
/* average duration (ms) of the solver's SpMV (y = A x, current matrix) over
 * `reps` launches on the context's stream; precision selects the FP64 or the
 * FP32 copy of the matrix.  Measurement aid, no counterpart in the reference. */
int mhfem_benchmark_spmv(mhfem_ctx *ctx, int precision, int reps, double *avg_ms);
/* AMG inspection (no counterpart in the reference): levels of the hierarchy
 * and rebuilds so far; rows, nonzeros and the aggregate of every row
 * (rows entries, NULL allowed, none on the coarsest level) of one level; one
 * V-cycle x = M^{-1} b on host vectors of the system's length. */
int mhfem_amg_info(const mhfem_ctx *ctx, int *levels, int *rebuilds);
int mhfem_amg_level(const mhfem_ctx *ctx, int level, int *rows, long *nnz, int32_t *aggregates);
int mhfem_amg_apply(mhfem_ctx *ctx, const double *b, double *x);
/* ILU-ISAI inspection (no counterpart in the reference): nonzeros of M_L and
 * M_U and rebuilds so far; M_L (upper = 0) or M_U (upper = 1) in CSR
 * (rowptr: rows + 1 entries, col/val: the nonzeros; NULL outputs skipped). */
int mhfem_isai_info(const mhfem_ctx *ctx, long *nnz_lower, long *nnz_upper, int *rebuilds);
int mhfem_isai_matrix(mhfem_ctx *ctx, int upper, int32_t *rowptr, int32_t *col, double *val);
int mhfem_get_timers(const mhfem_ctx *ctx, mhfem_timers *out);
int mhfem_reset_timers(mhfem_ctx *ctx);
int mhfem_set_timing(mhfem_ctx *ctx, int level); /* 0 off (default), 1 HIP-event phase timers,
                                                    2 phase timers + sampled SpMV launch timing */
int mhfem_synchronize(mhfem_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
