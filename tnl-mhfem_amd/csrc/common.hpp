// common.hpp -- context, device views and helpers shared by the MHFEM HIP
// translation units.  Everything here is internal to libmhfem_amd.so; the
// public surface is include/mhfem.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <algorithm>
#include <vector>

#include "../../include/mhfem.h"

namespace mhfem {

// ---------------------------------------------------------------------------
// error handling
struct Error {
    int code;
    std::string msg;
};

// sets the message mhfem_last_error(NULL) returns (calls without a context)
void set_global_error(const std::string &msg);

// device scalar slots of the Krylov solver (per segment)
enum Slot {
    S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_BETA, S_RAST_SQ, S_BNORM, S_RESID, S_TOL, S_IT,
    S_DONE, S_RESTART, S_MAXIT, S_D0, S_D1, S_D2, S_D3, S_TOLABS, S_DELTA, S_BFIX, S_BREAK, NSLOTS = 32
};
// a segment's solve has stopped (converged, failed or waiting for a restart):
// kernels queued behind it return at once
__device__ __forceinline__ bool halted(const double *sc) { return sc[S_DONE] != 0.0 || sc[S_RESTART] != 0.0; }

#define MHFEM_HIP(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::mhfem::Error{MHFEM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// ---------------------------------------------------------------------------
// A model coefficient: value(idx) = ptr[idx & mask]; mask = 0 for a uniform
// coefficient (UniformCoefficient.h:7-53), mask = ~0 for a full array.
struct Coef {
    const double *ptr;
    long mask;
    __device__ __forceinline__ double operator[](long idx) const { return ptr[idx & mask]; }
};

// Block SELL-64 storage of the hybridised system.  The unknowns of face E
// are the rows E*n+i, i < n (Solver.hpp:560-569); row (E,i) couples to the
// unknowns (F,j) of every face F of the cells of E, so the matrix is made of
// n x n blocks on a face-to-face pattern.  Faces are grouped in slices of 64
// (one wavefront, one lane per face).  Off-diagonal block k of face E sits at
// block index q = slice_ptr[E/64] + k*64 + E%64 with
//   col[q]                 = F (the column face),
//   val[q*n*n + i*n + j]   = A[(E,i), (F,j)],
// so a wavefront reading block k of its 64 faces touches one contiguous
// segment.  The assembled rows are scaled to a unit diagonal
// (LinearSystem.h:196-201, BoundaryConditions.hpp:200-203, :375-377): the
// block on the face's own column is stored without its diagonal, its n(n-1)
// other entries at val[diag_base + E*n*(n-1) + i*(n-1) + (j < i ? j : j-1)],
// and A x = x + (stored part) x.  Padding blocks have col = the face itself
// and zero values; so do the entries of a face's block rows that are not in
// that row's pattern (rows of one face may differ, e.g. a Dirichlet row).
template <typename T>
struct SellView {
    int faces;
    const int64_t *slice_ptr;  // in blocks
    const int *slice_width;    // blocks per face in the slice
    const int *col;            // column face of each block
    T *val;
    int64_t diag_base;
};

__host__ __device__ inline int64_t sell_block(const int64_t *slice_ptr, int face, int k)
{
    return slice_ptr[face >> 6] + (int64_t)k * 64 + (face & 63);
}

// position of entry (i, j) of the diagonal block of `face` (i != j)
__host__ __device__ inline int64_t sell_diag(int64_t diag_base, int n, int face, int i, int j)
{
    return diag_base + (int64_t)face * n * (n - 1) + i * (n - 1) + (j < i ? j : j - 1);
}

// M consecutive values moved as one vector access (16-byte aligned where the
// size allows; the storage offsets keep that alignment)
template <typename T, int M>
struct alignas((sizeof(T) * M) % 16 == 0 ? 16 : ((sizeof(T) * M) % 8 == 0 ? 8 : sizeof(T))) Pack {
    T v[M];
};
template <typename T, int M>
__device__ __forceinline__ Pack<T, M> ld(const T *p)
{
    return *reinterpret_cast<const Pack<T, M> *>(p);
}
// The same access with a non-temporal (streaming) hint: for data read once
// per sweep (matrix values and column indices), so that it does not push
// reused data (the gathered vector) out of the caches.
template <typename T, int M>
__device__ __forceinline__ Pack<T, M> ld_stream(const T *p)
{
    constexpr int B = (int)sizeof(T) * M;
    if constexpr (B % 16 == 0) {
        typedef unsigned int V __attribute__((ext_vector_type(4)));
        struct Raw {
            V c[B / 16];
        } r;
        const V *q = reinterpret_cast<const V *>(p);
#pragma unroll
        for (int c = 0; c < B / 16; c++) r.c[c] = __builtin_nontemporal_load(q + c);
        return __builtin_bit_cast(Pack<T, M>, r);
    } else if constexpr (B % 8 == 0) {
        typedef unsigned int V __attribute__((ext_vector_type(2)));
        struct Raw {
            V c[B / 8];
        } r;
        const V *q = reinterpret_cast<const V *>(p);
#pragma unroll
        for (int c = 0; c < B / 8; c++) r.c[c] = __builtin_nontemporal_load(q + c);
        return __builtin_bit_cast(Pack<T, M>, r);
    } else {
        Pack<T, M> r;
#pragma unroll
        for (int c = 0; c < M; c++) r.v[c] = __builtin_nontemporal_load(p + c);
        return r;
    }
}

template <typename T, int M>
__device__ __forceinline__ void st(T *p, const Pack<T, M> &v)
{
    *reinterpret_cast<Pack<T, M> *>(p) = v;
}

// Time function of a boundary model (mhfem_time_function), as the device
// reads it: a boundary value of species i is the stored value times the
// function at the step's new time (Solver.hpp:948 passes time + tau).
struct TimeFn {
    int kind;    // MHFEM_TIME_*, or -1: stationary (factor 1)
    int points;  // MHFEM_TIME_TABLE
    double offset, amplitude, omega, phase;
    const double *times, *values;  // MHFEM_TIME_TABLE, device
};

__device__ inline double time_factor(const TimeFn &f, double t)
{
    if (f.kind == MHFEM_TIME_CONSTANT) return f.offset;
    if (f.kind == MHFEM_TIME_SINE) return f.offset + f.amplitude * sin(f.omega * t + f.phase);
    if (f.kind == MHFEM_TIME_TABLE) {
        const int last = f.points - 1;
        if (!(t > f.times[0])) return f.values[0];
        if (!(t < f.times[last])) return f.values[last];
        int lo = 0, hi = last;  // times[lo] < t < times[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (f.times[mid] <= t) lo = mid;
            else hi = mid;
        }
        const double w = (t - f.times[lo]) / (f.times[hi] - f.times[lo]);
        return f.values[lo] + w * (f.values[hi] - f.values[lo]);
    }
    return 1.0;
}

// All device pointers of the scheme, passed to kernels by value.
struct SchemeView {
    int C, F, dim, fpc, S;
    int lumping, advection, do_mob;
    // topology
    const int *cell_faces;      // C*FPC, K-major
    const int *cell_vertices;   // C*NVC (NVC = FPC for simplices, 2^dim for boxes)
    const double *points;       // nv*dim
    const int *face_cells;      // F*2
    const uint64_t *face_slots; // per face: merged column order (see build_pattern)
    const uint32_t *face_blocks; // per face: block index of each cell's local faces (build_face_blocks)
    const int *bnd_faces;       // the boundary faces, ascending
    int nbnd;
    const double *cell_meas, *face_meas;
    // coefficients
    Coef N, u, m, D, w, a, r, f;
    // scheme arrays
    double *Z_iF, *Z_iK, *v, *m_upw, *Z_upw;
    const double *minv;         // M^{-1} per cell, packed planes (fixed geometry)
    const double *minv_rows;    // M^{-1} per cell, row-expanded (fixed geometry)
    double *cell_rec;           // per-cell assembly inputs written by k_cell
    // boundary conditions
    const uint8_t *tags;
    const double *bc_values, *bc_dirichlet, *bc_mobility;
    // boundary model: 2n time functions (Dirichlet of species i at [i],
    // Neumann at [n + i]), nullptr for the stationary model
    const TimeFn *bc_fn;
    const double *clock;  // [0]: the step's new time (time + tau), device
    // linear system
    SellView<double> A;
    double *rhs;
};

// ---------------------------------------------------------------------------
// device buffer
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        release();
        if (count) MHFEM_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

struct Krylov;  // krylov.hip
struct Ilu0;    // krylov.hip
struct Guess;   // guess.hip
struct Amg;     // amg.hip

struct Ctx {
    int dim = 0, fpc = 0, S = 0, n = 0;
    int box = 0, nvc = 0;  // box cells (edges, rectangles) or simplices; vertices per cell
    int nv = 0, C = 0, F = 0, nbnd = 0;
    int lumping = 0, advection = 0, do_mob = 1;
    int device = 0;  // the HIP device current at creation
    hipStream_t stream = nullptr;
    // host topology; position of each face (the lowest corner of its first
    // cell's bounding box, 3 floats; empty for batched instances), for the
    // ILU(0) solves' chunks
    std::vector<int> h_cell_faces, h_face_cells;
    std::vector<float> h_face_pos;
    // segments (batched ensembles): cells, faces and rows of each instance
    std::vector<int> h_seg_cells, h_seg_faces, h_seg_rows;
    // device topology
    DBuf<int> cell_faces, cell_vertices, face_cells, bnd_faces;
    DBuf<uint64_t> face_slots;
    DBuf<uint32_t> face_blocks;
    DBuf<double> points, cell_meas, face_meas;
    DBuf<double> minv, minv_rows;  // inverse geometric mass matrices: packed planes, row-expanded
    DBuf<double> cell_rec;  // per-cell records of the assembly inputs (k_cell)
    // model arrays (index by MHFEM_* ids); uniform coefficients hold 1 value
    DBuf<double> arr[MHFEM_NUM_ARRAYS];
    bool uniform[MHFEM_NUM_ARRAYS] = {};
    DBuf<uint8_t> tags;
    std::vector<uint8_t> h_tags;
    bool bc_set = false;
    // boundary model (mhfem_set_boundary_time_function): time functions of
    // the Dirichlet (slot i) and Neumann (slot n + i) values, their tables
    TimeFn h_bc_fn[6] = {};
    DBuf<TimeFn> bc_fn;
    DBuf<double> bc_tab[6];
    bool bc_timed = false;
    DBuf<double> clock;  // the step's new time, written before each assembly
    // linear system (block SELL-64 storage + CSR pattern on the host)
    int rows = 0;
    long nnz = 0;
    std::vector<int> h_rowptr, h_col;
    std::vector<int64_t> h_slice_ptr;
    std::vector<int> h_slice_width;
    std::vector<int64_t> h_bptr;  // per face: its stored blocks' column faces h_bface[h_bptr[E] ..]
    std::vector<int> h_bface;
    DBuf<int64_t> slice_ptr;
    DBuf<int> slice_width, sell_col, rowptr, csr_col;
    DBuf<int64_t> csr_to_sell;  // CSR entry -> position in sell_val, -1 for the unit diagonal
    DBuf<double> sell_val;
    long sell_size = 0;         // values in sell_val (blocks * n * n + F * n * (n - 1))
    int64_t sell_blocks = 0;    // block slots, padding included
    int64_t diag_base = 0;
    bool system_ready = false;
    bool unit_diagonal = true;  // the assembler writes 1 on every diagonal
    mhfem_solver_config cfg{};
    Krylov *krylov = nullptr;
    Guess *guess = nullptr;
    bool guess_on_solve = false;  // the learned guess runs in mhfem_solve
    // timers
    bool timing = false;
    int timing_level = 0;
    hipEvent_t ev[16] = {};
    // sampled SpMV launch timing (level 2): event pairs awaiting readout
    std::vector<hipEvent_t> kt_ev;
    size_t kt_used = 0;
    long kt_count = 0;
    std::vector<float> kt_samples;  // durations read back (ms)
    mhfem_timers timers{};
    std::string err;

    long array_size(int which) const;
    SchemeView view();
};

// Sampled kernel timing: kernel_timing_begin returns true when this launch
// is to be timed (then it has recorded the opening event), and
// kernel_timing_end records the closing one; the durations are read out by
// kernel_timing_flush.  Every 9th launch is timed: an odd period, so that
// the samples alternate between the two products of a BiCGStab iteration.
inline void kernel_timing_flush(Ctx &c)
{
    for (size_t q = 0; q + 1 < c.kt_used; q += 2) {
        MHFEM_HIP(hipEventSynchronize(c.kt_ev[q + 1]));
        float ms = 0;
        MHFEM_HIP(hipEventElapsedTime(&ms, c.kt_ev[q], c.kt_ev[q + 1]));
        c.kt_samples.push_back(ms);
    }
    c.kt_used = 0;
}
// Average of the sampled launches that did the product: a launch queued
// after its segments converged returns at once (a few microseconds against
// tens), so samples below a quarter of the median are left out.
inline void kernel_timing_summary(const Ctx &c, double &ms_sum, long long &count)
{
    ms_sum = 0.0;
    count = 0;
    if (c.kt_samples.empty()) return;
    std::vector<float> v(c.kt_samples);
    std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
    const float cut = 0.25f * v[v.size() / 2];
    for (float ms : c.kt_samples)
        if (ms >= cut) {
            ms_sum += ms;
            count++;
        }
}
inline bool kernel_timing_begin(Ctx &c)
{
    if (!c.timing || c.timing_level < 2 || c.kt_count++ % 9 != 0) return false;
    if (c.kt_used + 2 > c.kt_ev.size()) {
        if (c.kt_ev.size() < 512) {
            for (int q = 0; q < 64; q++) {
                hipEvent_t e;
                MHFEM_HIP(hipEventCreate(&e));
                c.kt_ev.push_back(e);
            }
        } else {
            kernel_timing_flush(c);
        }
    }
    MHFEM_HIP(hipEventRecord(c.kt_ev[c.kt_used++], c.stream));
    return true;
}
inline void kernel_timing_end(Ctx &c) { MHFEM_HIP(hipEventRecord(c.kt_ev[c.kt_used++], c.stream)); }

// scheme.hip
void launch_face_average(Ctx &c);
// the time the boundary model is evaluated at (clock[0] = t), on the stream
void launch_set_clock(Ctx &c, double t);
void launch_pre_iterate(Ctx &c, double time, double tau);
void launch_assemble(Ctx &c, double time, double tau);
void launch_post_iterate(Ctx &c);
void launch_geometry(Ctx &c);
void launch_fill_b(Ctx &c);
void launch_fill_R(Ctx &c);
void launch_zero_double(Ctx &c, double *p, size_t n);

// krylov.hip
Krylov *krylov_create(Ctx &c);
void krylov_destroy(Krylov *k);
// solve A x = b (b, x device arrays of length rows); returns stats
void krylov_solve(Ctx &c, const double *b, double *x, mhfem_solve_stats *st);
void krylov_bench_spmv(Ctx &c, int precision, int reps, double *avg_ms);
int krylov_autotune(Ctx &c, double *best_ms);
void krylov_segment_stats(Ctx &c, int *iterations, double *residues, int count);
void krylov_graph_step(Ctx &c, double time, double tau, mhfem_solve_stats *st);
void spmv64(Ctx &c, const double *x, double *y);
// learned-guess history: HIST_SLOTS FP64 slots per row, kept together
constexpr int HIST_SLOTS = 16;
void spmm_hist(Ctx &c, const double *H, double *AH, int slots);
void residual64(Ctx &c, const double *x, const double *b, double *r);
// the same, skipped once the single segment's solve has halted (sc: its flags, or nullptr)
void residual64_halt(Ctx &c, const double *x, const double *b, double *r, const double *sc);
double dot64(Ctx &c, const double *x, const double *y);
// the assembled matrix's values in CSR order (unit diagonal included) into dst[nnz]
void launch_csr_gather(Ctx &c, double *dst);

inline bool is_isai(int prec) { return prec == MHFEM_PREC_ILU_ISAI || prec == MHFEM_PREC_PARILU_ISAI; }
Amg *krylov_amg(const Krylov &k);  // the AMG preconditioner of a solver, or nullptr
// the ILU-ISAI preconditioner's sizes and rebuild count (false: none); copy of
// M_L (upper = 0) or M_U (upper = 1) in CSR, NULL outputs skipped
bool krylov_isai_info(const Krylov &k, long *nnz_lower, long *nnz_upper, int *rebuilds);
void krylov_isai_export(Ctx &c, int upper, int32_t *rowptr, int32_t *col, double *val);
// amg.hip: algebraic multigrid preconditioner (FP64)
Amg *amg_create();
void amg_destroy(Amg *a);
void amg_update(Ctx &c, Amg &a);
void amg_apply(Ctx &c, Amg &a, const double *x, double *y, const double *halt_flags);
void amg_after_solve(Amg &a, long long iterations);
int amg_num_levels(const Amg &a);
int amg_rebuilds(const Amg &a);
void amg_level_info(const Amg &a, int level, int *rows, long *nnz, int32_t *aggregates);

// guess.hip
Guess *guess_create(Ctx &c);
void guess_destroy(Guess *g);
void guess_configure(Ctx &c, Guess &g, const mhfem_guess_config &cfg);
void guess_before_solve(Ctx &c, Guess &g, const double *b, double *x);
void guess_after_solve(Ctx &c, Guess &g, const double *x);
void guess_forward(Ctx &c, Guess &g, const uint16_t *X, long rows, float *y);
void guess_gradient(Ctx &c, Guess &g, const uint16_t *X, const float *t, long rows, float *grad);
void guess_get_weights(Ctx &c, Guess &g, float *out);
void guess_set_weights(Ctx &c, Guess &g, const float *in);
int guess_num_params();
mhfem_guess_stats guess_last_stats(const Guess &g);

}  // namespace mhfem
