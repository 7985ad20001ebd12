// krylov.hip -- linear solvers for the hybridised MHFEM system on gfx950.
//
// * SELL-64 SpMV: one lane per row, a wavefront covers one 64-row slice, so
//   every matrix load of a wavefront is one contiguous 512-byte segment.
// * Segmented BiCGStab (left preconditioned, van der Vorst 1992).  The
//   unknowns are split into segments -- one for an ordinary solve, one per
//   instance for a batched ensemble (block-diagonal system).  Every work-group
//   owns a contiguous row range inside one segment; each segment carries its
//   own scalar recurrence (alpha, omega, beta, rho, residue, iteration count)
//   and DONE/RESTART flags in device memory.  Dot products reduce wavefront ->
//   work-group -> one partial per work-group, and a finalize kernel (one
//   work-group per segment) sums the partials in a fixed order and evaluates
//   the recurrence, so the iteration is deterministic and the host only polls
//   the flags every `check_every` iterations.  Work queued after a segment
//   converged sees its DONE flag and returns; a whole step can therefore be
//   captured once as a HIP graph and replayed.
// * GMRES(m), left preconditioned, classical Gram-Schmidt with one
//   re-orthogonalisation pass; Givens rotations on the host (one segment).
// * Jacobi (diagonal) and ILU(0) (Saad Alg. 10.4, exact, level-scheduled:
//   rows of one dependency level are processed by one launch).
// * FP32 inner BiCGStab inside FP64 iterative refinement (one segment).
// * An autotuner picks the rows per work-group of the fused iteration kernels
//   by timing them on the assembled pattern.
//
// Solver call sites in the reference: Solver.hpp:1230-1265 (TNL LinearSolver +
// Preconditioner chosen by name, configSetup.h:31-62).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "common.hpp"

namespace mhfem {

constexpr int KB = 256;
constexpr int MAXG = 2048;

enum Fin { FIN_NONE = 0, FIN_ALPHA, FIN_OMEGA, FIN_UPDATE, FIN_BNORM, FIN_INIT, FIN_REPLACE,
           FIN_ALPHA_MERGED, FIN_OMEGA_MERGED, FIN_REPLACE_MERGED };
// which segments a kernel works on
enum Mode { M_ALL = 0, M_ACTIVE = 1, M_RESTART = 2 };

// Segmented work decomposition.  Work-group b covers rows
// [blk_row[b], blk_row[b+1]) of segment blk_seg[b]; segment s owns the
// work-groups [seg_blk[s], seg_blk[s+1]).
struct Seg {
    int nseg, nblk;
    const int *blk_seg, *blk_row, *seg_blk;
    double *sc;        // [nseg][NSLOTS]
    double *partials;  // [ND][nblk]
    // a launch over nact work-groups that take the blocks active[0..nact);
    // nullptr: nact == nblk work-groups take every block
    const int *active;
    int nact;
};


__device__ __forceinline__ bool seg_selected(const double *sc, int mode)
{
    if (mode == M_ACTIVE) return !halted(sc);
    if (mode == M_RESTART) return sc[S_RESTART] != 0.0;
    return true;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ void finalize(double *sc, int fin)
{
    const double eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
    switch (fin) {
    case FIN_ALPHA: {
        // breakdown (r*, A p) = 0: the segment restarts from its current x
        // (the rest of the iteration skips it); counted as an iteration so
        // that a segment that keeps breaking down still ends at max-iterations
        const double a = sc[S_RHO] / sc[S_D0];
        if (!(fabs(sc[S_D0]) > 0.0) || !isfinite(a)) {
            sc[S_RESTART] = 1;
            sc[S_IT] += 1;
            if (sc[S_IT] >= sc[S_MAXIT]) sc[S_DONE] = 1;
        } else {
            sc[S_ALPHA] = a;
        }
        break;
    }
    case FIN_OMEGA: sc[S_OMEGA] = sc[S_D1] > 0 ? sc[S_D0] / sc[S_D1] : 0.0; break;
    case FIN_UPDATE: {
        sc[S_RHO_OLD] = sc[S_RHO];
        sc[S_RHO] = sc[S_D0];
        sc[S_IT] += 1;
        if (!(fabs(sc[S_RHO]) >= eps2 * sc[S_RAST_SQ]) || sc[S_OMEGA] == 0.0) {
            // rho breakdown, or omega = 0 (the next beta would divide by it):
            // the host recomputes the residual and restarts the recurrence
            sc[S_RESTART] = 1;
        } else {
            sc[S_BETA] = (sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]);
            const double res = sqrt(sc[S_D1]) / sc[S_BNORM];
            sc[S_RESID] = res;
            if (!(res >= sc[S_TOL]) || sc[S_IT] >= sc[S_MAXIT]) sc[S_DONE] = 1;
        }
        break;
    }
    case FIN_BNORM: {
        const double bn = sqrt(sc[S_D0]);
        sc[S_BNORM] = bn == 0.0 ? 1.0 : bn;
        break;
    }
    case FIN_INIT: {
        sc[S_BREAK] = 0;
        sc[S_RHO] = sc[S_D0];
        sc[S_RAST_SQ] = sc[S_D0];
        const double res = sqrt(sc[S_D0]) / sc[S_BNORM];
        sc[S_RESID] = res;
        sc[S_RESTART] = 0;
        sc[S_DONE] = (!(res >= sc[S_TOL]) || sc[S_IT] >= sc[S_MAXIT]) ? 1 : 0;
        break;
    }
    case FIN_REPLACE: {
        // reliable update: the recursive residual was replaced by the true
        // one (D0 = (r, r), D1 = (r, r_ast)); the recurrence continues with
        // the same search direction, so only rho, beta and the reference
        // norm of the next stage change
        const double rn = sqrt(sc[S_D0]);
        sc[S_RHO] = sc[S_D1];
        sc[S_BETA] = (sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]);
        sc[S_BNORM] = rn;
        sc[S_TOL] = rn > 0 ? fmax(sc[S_DELTA], sc[S_TOLABS] / rn) : 1.0;
        sc[S_RESID] = 1.0;
        sc[S_DONE] = rn == 0.0 ? 1 : 0;
        // breakdown of the continued recurrence: start afresh from the
        // replaced residual
        sc[S_RESTART] = (rn > 0.0 && !(fabs(sc[S_RHO]) >= eps2 * sc[S_RAST_SQ])) ? 1 : 0;
        break;
    }
    // The merged iteration (FP32 inner solves): one reduction fewer per
    // iteration.  The convergence test on (r, r), which the x/r/p update of
    // the previous iteration summed (D1), comes with alpha (D0 = (r*, A p)).
    case FIN_ALPHA_MERGED: {
        const double res = sqrt(sc[S_D1]) / sc[S_BNORM];
        sc[S_RESID] = res;
        if (!(res >= sc[S_TOL]) || sc[S_IT] >= sc[S_MAXIT]) {
            sc[S_DONE] = 1;
            break;
        }
        if (sc[S_BREAK] != 0.0) {
            // the previous iteration broke down: restart from its x
            sc[S_BREAK] = 0;
            sc[S_RESTART] = 1;
            break;
        }
        const double a = sc[S_RHO] / sc[S_D0];
        if (!(fabs(sc[S_D0]) > 0.0) || !isfinite(a)) {
            sc[S_RESTART] = 1;
            sc[S_IT] += 1;
            if (sc[S_IT] >= sc[S_MAXIT]) sc[S_DONE] = 1;
        } else {
            sc[S_ALPHA] = a;
        }
        break;
    }
    // omega = (As, s) / (As, As); the next rho = (r, r*) with r = s - omega As
    // is (s, r*) - omega (As, r*) (D3, D2), so beta is known before r is
    // formed and p is updated in the same pass as x and r
    case FIN_OMEGA_MERGED: {
        const double om = sc[S_D1] > 0 ? sc[S_D0] / sc[S_D1] : 0.0;
        const double rho = sc[S_D3] - om * sc[S_D2];
        sc[S_OMEGA] = om;
        sc[S_IT] += 1;
        sc[S_RHO_OLD] = sc[S_RHO];
        sc[S_RHO] = rho;
        if (!(fabs(rho) >= eps2 * sc[S_RAST_SQ]) || om == 0.0) {
            // breakdown: x and r are still updated (p = r), then the next
            // alpha step tests convergence and otherwise asks the host to
            // recompute the residual and restart
            sc[S_BETA] = 0;
            sc[S_BREAK] = 1;
        } else {
            sc[S_BETA] = (rho / sc[S_RHO_OLD]) * (sc[S_ALPHA] / om);
        }
        break;
    }
    // reliable update in the merged iteration (D0 = D1 = (r, r), D2 = (r, r*)
    // of the replaced residual).  p was already advanced with the recursive
    // residual: p = r_rec + beta d, d = p_old - omega A p_old.  The
    // continued direction is r + beta' d = r + (beta' / beta) (p - r_rec);
    // S_BFIX is that ratio (0: the direction restarts from r)
    case FIN_REPLACE_MERGED: {
        const double rn = sqrt(sc[S_D0]);
        const double beta_old = sc[S_BETA];
        sc[S_RHO] = sc[S_D2];
        sc[S_BETA] = (sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]);
        const double fix = sc[S_BETA] / beta_old;
        sc[S_BFIX] = isfinite(fix) ? fix : 0.0;
        sc[S_BREAK] = 0;
        sc[S_BNORM] = rn;
        sc[S_TOL] = rn > 0 ? fmax(sc[S_DELTA], sc[S_TOLABS] / rn) : 1.0;
        sc[S_RESID] = 1.0;
        sc[S_DONE] = rn == 0.0 ? 1 : 0;
        sc[S_RESTART] = (rn > 0.0 && !(fabs(sc[S_RHO]) >= eps2 * sc[S_RAST_SQ])) ? 1 : 0;
        break;
    }
    default: break;
    }
}

// One partial sum per work-group (wavefront shuffles, then LDS).  The
// partials are summed by k_finalize, launched next on the same stream: the
// kernel boundary orders the writes, so no work-group issues a device-scope
// fence (on gfx950 that writes back the L2 of its XCD, and once per
// work-group it costs more than the kernel's own streaming).
// Work-group -> work item.  Work-groups are dealt to the 8 XCDs of the chip
// in turn (work-group b runs on XCD b % 8); renumbering them so that XCD k
// takes the k-th contiguous eighth of the work items keeps the faces that
// share column faces -- and therefore their x gathers -- in one L2.  Any
// bijection is correct; this one only changes where the work runs.
__device__ __forceinline__ int xcd_block(int nblk)
{
    const int b = blockIdx.x, k = b & 7, j = b >> 3;
    const int q = nblk >> 3, r = nblk & 7;
    return k * q + (k < r ? k : r) + j;
}

template <int ND, int OFF = 0, int BS = KB>
__device__ __forceinline__ void block_partial(double (&v)[ND], const Seg &sg, int blk)
{
    __shared__ double sm[ND][BS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const double w = wave_sum(v[d]);
        if (lane == 0) sm[d][wid] = w;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
            double s = sm[d][0];
#pragma unroll
            for (int q = 1; q < BS / 64; q++) s += sm[d][q];
            sg.partials[(long)(d + OFF) * sg.nblk + blk] = s;
        }
    }
}

// Sum the partials of each selected segment into sc[S_D0 + d] and evaluate
// the scalar recurrence `fin`.  One work-group per segment.  Thread tid
// sums the partials b0 + tid + m KB of each row in four chains, m going to
// chain m mod 4 while a whole group of four is left and to chain 0 after
// (a fixed order: the sum is the same for a given decomposition).  The
// loads of eight values of every row are issued at once, clamped to the
// segment's last partial (adding zero), with no branch between them, so a
// single-segment system with thousands of work-groups reads its partials
// in a few rounds of memory latency.
template <int ND>
__global__ void __launch_bounds__(KB) k_finalize(Seg sg, int fin, int mode)
{
    const int s = blockIdx.x;
    double *sc = sg.sc + (long)s * NSLOTS;
    if (!seg_selected(sc, mode)) return;
    __shared__ double sm[ND][KB / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int b0 = sg.seg_blk[s], b1 = sg.seg_blk[s + 1];
    const int first = b0 + tid;
    const int M = first < b1 ? (b1 - first + KB - 1) / KB : 0;  // partials of this thread
    const int full = M & ~3;
    double a[ND][4];
#pragma unroll
    for (int d = 0; d < ND; d++) a[d][0] = a[d][1] = a[d][2] = a[d][3] = 0.0;
    for (int q = 0; q < M; q += 8) {
        double v[ND][8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int i = min(first + (q + j) * KB, b1 - 1);
#pragma unroll
            for (int d = 0; d < ND; d++) v[d][j] = sg.partials[(long)d * sg.nblk + i];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int m = q + j;
#pragma unroll
            for (int d = 0; d < ND; d++) {
                a[d][j & 3] += m < full ? v[d][j] : 0.0;
                a[d][0] += m >= full && m < M ? v[d][j] : 0.0;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const double w = wave_sum((a[d][0] + a[d][1]) + (a[d][2] + a[d][3]));
        if (lane == 0) sm[d][wid] = w;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
            double v = sm[d][0];
#pragma unroll
            for (int q = 1; q < KB / 64; q++) v += sm[d][q];
            sc[S_D0 + d] = v;
        }
        finalize(sc, fin);
    }
}

// scalar state of every segment at the start of a solve (a kernel rather
// than a host copy, so that it can be captured in a graph)
__global__ void k_init_scalars(double *sc, int nseg, double tol, double maxit)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double *p = sc + (long)s * NSLOTS;
    for (int k = 0; k < NSLOTS; k++) p[k] = 0.0;
    p[S_TOL] = tol;
    p[S_MAXIT] = maxit;
}

// ---------------------------------------------------------------------------
// The N rows of face E: y = x_E + (diagonal block without its unit
// diagonal) x_E + sum over the stored blocks of B_k x_{col_k}.  One block
// costs one column index, one N*N vector load and one N-wide gather of x.
// The loads are issued in three waves -- the column indices of the face's
// first PF blocks, then those blocks and their gathers, then the products --
// so that a lane keeps PF blocks in flight instead of waiting on each index
// and gather in turn.  The slice width is uniform across a wavefront's 64
// faces, so the k < w guards do not diverge.  Measured at n = 2 on the 64^3
// Kuhn mesh: FP32 runs best with PF = 4 and streaming (non-temporal) matrix
// loads, which leave the caches to the gathered x (whole time step of the
// mixed-precision solve 36.9 ms; 37.9 ms at PF = 6, 38.7 ms at PF = 8,
// 39.6 ms at PF = 3); FP64 with PF = 8 and ordinary loads (isolated product
// 169 us; 183 us with streaming loads).  The FP32 product also loads the
// column indices of blocks PF and PF+1 up front, and those two blocks' values
// and gathers as one second batch after the first blocks' products (the
// empty asm keeps the compiler from hoisting them above those products):
// 58-64 VGPRs, 8 waves per SIMD, and one round of loads for the two instead
// of an index and then a gather for each in turn (3 % more time steps per
// second at n = 2 on 64^3; issuing the second batch with the first, 56-62
// VGPRs, was 2 % slower).
template <int N, typename T>
__device__ __forceinline__ Pack<T, N> sell_face(const SellView<T> &A, const T *__restrict__ x, int E,
                                                Pack<T, N> &xe_out)
{
    constexpr bool fp32 = sizeof(T) == 4;
    constexpr int PF = N <= 2 ? (fp32 ? 4 : 8) : 3;
    constexpr int P2 = fp32 && N <= 2 ? 2 : 0;
    const int64_t base = A.slice_ptr[E >> 6] + (E & 63);
    const int w = A.slice_width[E >> 6];
    int col[PF + P2];
#pragma unroll
    for (int k = 0; k < PF + P2; k++)
        if (k < w) {
            if constexpr (fp32) col[k] = __builtin_nontemporal_load(A.col + base + (int64_t)k * 64);
            else col[k] = A.col[base + (int64_t)k * 64];
        }
    const Pack<T, N> xe = ld<T, N>(x + (int64_t)E * N);
    xe_out = xe;
    Pack<T, N * N> b[PF];
    Pack<T, N> xf[PF];
#pragma unroll
    for (int k = 0; k < PF; k++)
        if (k < w) {
            if constexpr (fp32) b[k] = ld_stream<T, N * N>(A.val + (base + (int64_t)k * 64) * (N * N));
            else b[k] = ld<T, N * N>(A.val + (base + (int64_t)k * 64) * (N * N));
            xf[k] = ld<T, N>(x + (int64_t)col[k] * N);
        }
    Pack<T, N> y = xe;
    if constexpr (N > 1) {
        const Pack<T, N *(N - 1)> dg = ld_stream<T, N *(N - 1)>(A.val + A.diag_base + (int64_t)E * N * (N - 1));
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++)
                if (j != i) y.v[i] += dg.v[i * (N - 1) + (j < i ? j : j - 1)] * xe.v[j];
    }
#pragma unroll
    for (int k = 0; k < PF; k++)
        if (k < w) {
#pragma unroll
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int j = 0; j < N; j++) y.v[i] += b[k].v[i * N + j] * xf[k].v[j];
        }
    if constexpr (P2 > 0) {
        asm volatile("" ::: "memory");
        Pack<T, N * N> b2[P2 > 0 ? P2 : 1];
        Pack<T, N> x2[P2 > 0 ? P2 : 1];
#pragma unroll
        for (int k = 0; k < P2; k++)
            if (PF + k < w) {
                b2[k] = ld_stream<T, N * N>(A.val + (base + (int64_t)(PF + k) * 64) * (N * N));
                x2[k] = ld<T, N>(x + (int64_t)col[PF + k] * N);
            }
#pragma unroll
        for (int k = 0; k < P2; k++)
            if (PF + k < w) {
#pragma unroll
                for (int i = 0; i < N; i++)
#pragma unroll
                    for (int j = 0; j < N; j++) y.v[i] += b2[k].v[i * N + j] * x2[k].v[j];
            }
    }
    for (int k = PF + P2; k < w; k++) {
        const int64_t q = base + (int64_t)k * 64;
        const Pack<T, N * N> bk = ld<T, N * N>(A.val + q * (N * N));
        const Pack<T, N> xk = ld<T, N>(x + (int64_t)A.col[q] * N);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) y.v[i] += bk.v[i * N + j] * xk.v[j];
    }
    return y;
}

// the same without the face's own x values
template <int N, typename T>
__device__ __forceinline__ Pack<T, N> sell_face(const SellView<T> &A, const T *__restrict__ x, int E)
{
    Pack<T, N> xe;
    return sell_face<N>(A, x, E, xe);
}

// launch helper: f(std::integral_constant<int, n>) for the supported n
template <typename F>
static void with_n(int n, F &&f)
{
    switch (n) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    default: throw Error{MHFEM_E_ARGUMENT, "unsupported number of equations"};
    }
}

#define SEG_BLOCK_PROLOGUE(sg, mode)                                   \
    const int blk_ = (sg).active ? (sg).active[xcd_block((sg).nact)]   \
                                 : xcd_block((sg).nblk);               \
    const int seg_ = (sg).blk_seg[blk_];                               \
    double *sc_ = (sg).sc + (long)seg_ * NSLOTS;                       \
    if (!seg_selected(sc_, (mode))) return;                            \
    const int row0_ = (sg).blk_row[blk_], row1_ = (sg).blk_row[blk_ + 1]

// y = D (A x) (D = dinv or identity); ND fused dots: (y, z), (y, y), and
// for ND = 4 also (y, w), (z, w).  A work-group's row range holds whole
// faces (build_blocks).
// BS threads per work-group (256, or 512 for the FP32 product; sspmv()).
template <typename T, int N, bool DIAG, int ND, int BS>
__global__ void __launch_bounds__(BS) k_sspmv(SellView<T> A, Seg sg, int mode, const T *__restrict__ x,
                                              const T *__restrict__ dinv, T *__restrict__ y, const T *__restrict__ z,
                                              const T *__restrict__ w)
{
    SEG_BLOCK_PROLOGUE(sg, mode);
    double acc[ND > 0 ? ND : 1] = {};
    // the iteration's second product dots with its own input (z == x): the
    // face's x values loaded for the product serve as z
    const bool zx = z == x;
    for (int f = row0_ / N + threadIdx.x; f < row1_ / N; f += BS) {
        Pack<T, N> xe;
        Pack<T, N> v = sell_face<N>(A, x, f, xe);
        if constexpr (DIAG) {
            const Pack<T, N> d = ld<T, N>(dinv + (int64_t)f * N);
#pragma unroll
            for (int i = 0; i < N; i++) v.v[i] *= d.v[i];
        }
        st<T, N>(y + (int64_t)f * N, v);
        Pack<T, N> zz;
        if constexpr (ND >= 1) {
            zz = zx ? xe : ld<T, N>(z + (int64_t)f * N);
#pragma unroll
            for (int i = 0; i < N; i++) acc[0] += (double)v.v[i] * (double)zz.v[i];
        }
        if constexpr (ND >= 2) {
#pragma unroll
            for (int i = 0; i < N; i++) acc[1] += (double)v.v[i] * (double)v.v[i];
        }
        if constexpr (ND >= 4) {
            const Pack<T, N> ww = ld<T, N>(w + (int64_t)f * N);
#pragma unroll
            for (int i = 0; i < N; i++) {
                acc[2] += (double)v.v[i] * (double)ww.v[i];
                acc[3] += (double)zz.v[i] * (double)ww.v[i];
            }
        }
    }
    if constexpr (ND > 0) block_partial<ND, 0, BS>(acc, sg, blk_);
}

// y = b - A x on the selected segments
template <typename T, int N>
__global__ void __launch_bounds__(KB) k_sresidual(SellView<T> A, Seg sg, int mode, const T *__restrict__ x,
                                                  const T *__restrict__ b, T *__restrict__ y)
{
    SEG_BLOCK_PROLOGUE(sg, mode);
    for (int f = row0_ / N + threadIdx.x; f < row1_ / N; f += KB) {
        Pack<T, N> v = sell_face<N>(A, x, f);
        const Pack<T, N> bb = ld<T, N>(b + (int64_t)f * N);
#pragma unroll
        for (int i = 0; i < N; i++) v.v[i] = bb.v[i] - v.v[i];
        st<T, N>(y + (int64_t)f * N, v);
    }
}

template <typename T, int ND>
struct DotArgs {
    const T *x[ND];
    const T *y[ND];
};

template <typename T, int ND>
__global__ void __launch_bounds__(KB) k_sdots(Seg sg, int mode, DotArgs<T, ND> a)
{
    SEG_BLOCK_PROLOGUE(sg, mode);
    double acc[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) acc[d] = 0;
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB)
#pragma unroll
        for (int d = 0; d < ND; d++) acc[d] += (double)a.x[d][i] * (double)a.y[d][i];
    block_partial<ND>(acc, sg, blk_);
}

template <typename T>
__global__ void __launch_bounds__(KB) k_sscale(Seg sg, int mode, const T *__restrict__ x, const T *__restrict__ d,
                                               T *__restrict__ y)
{
    SEG_BLOCK_PROLOGUE(sg, mode);
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) y[i] = d ? x[i] * d[i] : x[i];
}

template <typename T>
__global__ void __launch_bounds__(KB) k_scopy2(Seg sg, int mode, const T *__restrict__ src, T *__restrict__ a,
                                               T *__restrict__ b)
{
    SEG_BLOCK_PROLOGUE(sg, mode);
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) {
        const T v = src[i];
        a[i] = v;
        if (b) b[i] = v;
    }
}

// s = r - alpha Ap
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_s(Seg sg, const T *__restrict__ r, const T *__restrict__ Ap,
                                               T *__restrict__ s)
{
    SEG_BLOCK_PROLOGUE(sg, M_ACTIVE);
    const T alpha = (T)sc_[S_ALPHA];
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) s[i] = r[i] - alpha * Ap[i];
}

// x += alpha p + omega s; r = s - omega As; dots (r, r_ast), (r, r)
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_xr(Seg sg, T *__restrict__ x, const T *__restrict__ p,
                                                const T *__restrict__ s, const T *__restrict__ As, T *__restrict__ r,
                                                const T *__restrict__ r_ast)
{
    SEG_BLOCK_PROLOGUE(sg, M_ACTIVE);
    const T alpha = (T)sc_[S_ALPHA], omega = (T)sc_[S_OMEGA];
    double acc[2] = {0, 0};
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) {
        const T si = s[i];
        x[i] += alpha * p[i] + omega * si;
        const T ri = si - omega * As[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)r_ast[i];
        acc[1] += (double)ri * (double)ri;
    }
    block_partial<2>(acc, sg, blk_);
}

// p = r + beta (p - omega Ap)
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_p(Seg sg, const T *__restrict__ r, T *__restrict__ p,
                                               const T *__restrict__ Ap)
{
    SEG_BLOCK_PROLOGUE(sg, M_ACTIVE);
    const T beta = (T)sc_[S_BETA], omega = (T)sc_[S_OMEGA];
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) p[i] = r[i] + beta * (p[i] - omega * Ap[i]);
}

// merged iteration: x += alpha p + omega s; r = s - omega As;
// p = r + beta (p - omega Ap); dot (r, r) into partial row 1 (read with
// the next alpha)
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_xrp(Seg sg, T *__restrict__ x, T *__restrict__ p,
                                                 const T *__restrict__ s, const T *__restrict__ As, T *__restrict__ r,
                                                 const T *__restrict__ Ap)
{
    SEG_BLOCK_PROLOGUE(sg, M_ACTIVE);
    const T alpha = (T)sc_[S_ALPHA], omega = (T)sc_[S_OMEGA], beta = (T)sc_[S_BETA];
    double acc[1] = {0};
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) {
        const T si = s[i], pi = p[i];
        x[i] += alpha * pi + omega * si;
        const T ri = si - omega * As[i];
        r[i] = ri;
        p[i] = ri + beta * (pi - omega * Ap[i]);
        acc[0] += (double)ri * (double)ri;
    }
    block_partial<1, 1>(acc, sg, blk_);
}

// merged iteration, after a reliable update: q = p - r (before r is
// replaced), then p = r + S_BFIX q (after)
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_dir(Seg sg, const T *__restrict__ p, const T *__restrict__ r,
                                                 T *__restrict__ q)
{
    SEG_BLOCK_PROLOGUE(sg, M_ALL);
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) q[i] = p[i] - r[i];
}
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_pfix(Seg sg, const T *__restrict__ r, T *__restrict__ p,
                                                  const T *__restrict__ q)
{
    SEG_BLOCK_PROLOGUE(sg, M_ACTIVE);
    const T f = (T)sc_[S_BFIX];
    for (int i = row0_ + threadIdx.x; i < row1_; i += KB) p[i] = r[i] + f * q[i];
}

// ---------------------------------------------------------------------------
// unsegmented helpers (no flags consulted)
template <typename T, int N>
__global__ void __launch_bounds__(KB) k_residual(SellView<T> A, const T *__restrict__ x, const T *__restrict__ b,
                                                 T *__restrict__ y)
{
    for (int f = blockIdx.x * KB + threadIdx.x; f < A.faces; f += gridDim.x * KB) {
        Pack<T, N> v = sell_face<N>(A, x, f);
        const Pack<T, N> bb = ld<T, N>(b + (int64_t)f * N);
#pragma unroll
        for (int i = 0; i < N; i++) v.v[i] = bb.v[i] - v.v[i];
        st<T, N>(y + (int64_t)f * N, v);
    }
}

// y = b - A x, skipped once the single segment has halted (the finest level
// of the AMG V-cycle)
template <int N>
__global__ void __launch_bounds__(KB) k_residual_halt(SellView<double> A, const double *__restrict__ x,
                                                      const double *__restrict__ b, double *__restrict__ y,
                                                      const double *sc)
{
    if (sc && halted(sc)) return;
    for (int f = blockIdx.x * KB + threadIdx.x; f < A.faces; f += gridDim.x * KB) {
        Pack<double, N> v = sell_face<N>(A, x, f);
        const Pack<double, N> bb = ld<double, N>(b + (int64_t)f * N);
#pragma unroll
        for (int i = 0; i < N; i++) v.v[i] = bb.v[i] - v.v[i];
        st<double, N>(y + (int64_t)f * N, v);
    }
}

template <typename T, int N>
__global__ void __launch_bounds__(KB) k_spmv_plain(SellView<T> A, const T *__restrict__ x, T *__restrict__ y)
{
    for (int f = blockIdx.x * KB + threadIdx.x; f < A.faces; f += gridDim.x * KB)
        st<T, N>(y + (int64_t)f * N, sell_face<N>(A, x, f));
}

// AH = A H for the stored increments of the learned guess.  H and AH keep
// the HIST_SLOTS slots of a row together (FP64, 128 bytes per row): the
// gather of a column face fetches the values of every slot from two cache
// lines, where one array per slot would touch a line per slot.  L lanes
// share a face, lane g taking the slots 4g..4g+3 (L = the groups of 4 slots
// in use): the L lanes read the face's blocks at the same addresses (one
// request per wavefront) and its rows' slots as one contiguous run, and a
// lane holds 4 slots of accumulators instead of 16 (about 60 VGPRs instead
// of 130, so 8 waves per SIMD instead of 3).  Each slot's sum runs in the
// same order as with one lane per face.
template <int N, int L>
__global__ void __launch_bounds__(KB) k_spmm_hist(SellView<double> A, const double *__restrict__ H,
                                                  double *__restrict__ AH)
{
    constexpr int PF = 8;
    const int64_t gid = (int64_t)xcd_block(gridDim.x) * KB + threadIdx.x;
    const int f = (int)(gid / L), g0 = (int)(gid % L);  // face, and this lane's group of 4 slots
    if (f >= A.faces) return;
    const int64_t base = A.slice_ptr[f >> 6] + (f & 63);
    const int w = A.slice_width[f >> 6];
    int col[PF];
#pragma unroll
    for (int k = 0; k < PF; k++)
        if (k < w) col[k] = A.col[base + (int64_t)k * 64];
    auto load_rows = [&](int face, double (&v)[N][4]) {
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double *p = H + ((int64_t)face * N + j) * HIST_SLOTS + 4 * g0;
            const Pack<double, 2> a = ld<double, 2>(p), b = ld<double, 2>(p + 2);
            v[j][0] = a.v[0];
            v[j][1] = a.v[1];
            v[j][2] = b.v[0];
            v[j][3] = b.v[1];
        }
    };
    double y[N][4];
    {
        // the unit diagonal and the stored entries of the diagonal block
        double xe[N][4];
        load_rows(f, xe);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int s = 0; s < 4; s++) y[i][s] = xe[i][s];
        if constexpr (N > 1) {
            const Pack<double, N *(N - 1)> dg = ld<double, N *(N - 1)>(A.val + A.diag_base + (int64_t)f * N * (N - 1));
#pragma unroll
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int j = 0; j < N; j++)
                    if (j != i) {
                        const double a = dg.v[i * (N - 1) + (j < i ? j : j - 1)];
#pragma unroll
                        for (int s = 0; s < 4; s++) y[i][s] += a * xe[j][s];
                    }
        }
    }
    auto block = [&](int64_t q, int c) {
        const Pack<double, N * N> b = ld<double, N * N>(A.val + q * (N * N));
        double xc[N][4];
        load_rows(c, xc);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++)
#pragma unroll
                for (int s = 0; s < 4; s++) y[i][s] += b.v[i * N + j] * xc[j][s];
    };
#pragma unroll
    for (int k = 0; k < PF; k++)
        if (k < w) block(base + (int64_t)k * 64, col[k]);
    for (int k = PF; k < w; k++) {
        const int64_t q = base + (int64_t)k * 64;
        block(q, A.col[q]);
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        double *o = AH + ((int64_t)f * N + i) * HIST_SLOTS + 4 * g0;
        Pack<double, 2> a, b;
        a.v[0] = y[i][0];
        a.v[1] = y[i][1];
        b.v[0] = y[i][2];
        b.v[1] = y[i][3];
        st<double, 2>(o, a);
        st<double, 2>(o + 2, b);
    }
}

// FP64 residual of the refinement loop, stored in FP32: r32 = (float)(b - A x),
// face f's rows at face iperm[f] of the FP32 iteration's order (identity if null)
template <int N>
__global__ void __launch_bounds__(KB) k_residual_mixed(SellView<double> A, const double *__restrict__ x,
                                                       const double *__restrict__ b, float *__restrict__ r32,
                                                       const int *__restrict__ iperm)
{
    for (int f = blockIdx.x * KB + threadIdx.x; f < A.faces; f += gridDim.x * KB) {
        const Pack<double, N> v = sell_face<N>(A, x, f);
        const Pack<double, N> bb = ld<double, N>(b + (int64_t)f * N);
        Pack<float, N> r;
#pragma unroll
        for (int i = 0; i < N; i++) r.v[i] = (float)(bb.v[i] - v.v[i]);
        st<float, N>(r32 + (int64_t)(iperm ? iperm[f] : f) * N, r);
    }
}

// The FP32 copy of the matrix in the gather-friendly face order: block b of
// the permuted layout takes natural block map[b] (-1: padding, zero); the
// within-face couplings of permuted face p are those of face perm[p].
template <int N, typename TO>
__global__ void __launch_bounds__(KB) k_convert_permuted(int64_t blocks, const int64_t *__restrict__ map,
                                                         const double *__restrict__ src, TO *__restrict__ dst)
{
    constexpr int NN = N * N;
    for (int64_t b = (int64_t)blockIdx.x * KB + threadIdx.x; b < blocks; b += (int64_t)gridDim.x * KB) {
        const int64_t q = map[b];
        Pack<TO, NN> o;
        if (q >= 0) {
            const Pack<double, NN> s = ld<double, NN>(src + q * NN);
#pragma unroll
            for (int t = 0; t < NN; t++) o.v[t] = (TO)s.v[t];
        } else {
#pragma unroll
            for (int t = 0; t < NN; t++) o.v[t] = TO(0);
        }
        st<TO, NN>(dst + b * NN, o);
    }
}
template <int N, typename TO>
__global__ void __launch_bounds__(KB) k_convert_permuted_diag(int faces, const int *__restrict__ perm,
                                                              const double *__restrict__ src, TO *__restrict__ dst)
{
    constexpr int D = N * (N - 1);
    for (int p = blockIdx.x * KB + threadIdx.x; p < faces; p += gridDim.x * KB) {
        const int64_t e = perm[p];
#pragma unroll
        for (int t = 0; t < D; t++) dst[(int64_t)p * D + t] = (TO)src[e * D + t];
    }
}

// the n values of every face into the permuted face order (FORWARD:
// out[p] = in[perm[p]]) or back (out[perm[p]] = in[p])
template <int N, bool FORWARD>
__global__ void __launch_bounds__(KB) k_permute_faces(int faces, const int *__restrict__ perm,
                                                      const double *__restrict__ in, double *__restrict__ out)
{
    for (int p = blockIdx.x * KB + threadIdx.x; p < faces; p += gridDim.x * KB) {
        const int64_t e = perm[p];
        if (FORWARD) st<double, N>(out + (int64_t)p * N, ld<double, N>(in + e * N));
        else st<double, N>(out + e * N, ld<double, N>(in + (int64_t)p * N));
    }
}

template <typename TO, typename TI>
__global__ void __launch_bounds__(KB) k_convert(long n, const TI *__restrict__ x, TO *__restrict__ y)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) y[i] = (TO)x[i];
}

// x64 += (double) d32; d32 = 0
__global__ void __launch_bounds__(KB) k_add_correction(int n, double *__restrict__ x, float *__restrict__ d)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        x[i] += (double)d[i];
        d[i] = 0.0f;
    }
}

// the same with the FP32 iteration's face order: x[f] += d32[iperm[f]]
template <int N>
__global__ void __launch_bounds__(KB) k_add_correction_perm(int faces, double *__restrict__ x,
                                                            float *__restrict__ d, const int *__restrict__ iperm)
{
    for (int f = blockIdx.x * KB + threadIdx.x; f < faces; f += gridDim.x * KB) {
        const int64_t q = (int64_t)iperm[f] * N;
#pragma unroll
        for (int i = 0; i < N; i++) {
            x[(int64_t)f * N + i] += (double)d[q + i];
            d[q + i] = 0.0f;
        }
    }
}

__global__ void k_set_slots(double *sc, double tolabs, double delta)
{
    if (threadIdx.x == 0) {
        sc[S_TOLABS] = tolabs;
        sc[S_DELTA] = delta;
    }
}

__global__ void k_set_slot(double *sc, int slot, double v)
{
    if (threadIdx.x == 0) sc[slot] = v;
}

template <typename T>
__global__ void __launch_bounds__(KB) k_fill_t(long n, T *__restrict__ x, T v)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) x[i] = v;
}

// ---------------------------------------------------------------------------
// ILU(0), level-scheduled.  Factors are kept in CSR order (natural rows).
// CSR-ordered copy of the matrix; map < 0 marks the implicit unit diagonal
template <typename T>
__global__ void __launch_bounds__(KB) k_gather(long n, const T *__restrict__ src, const int64_t *__restrict__ map,
                                               T *__restrict__ dst)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) {
        const int64_t q = map[i];
        dst[i] = q < 0 ? (T)1 : src[q];
    }
}

// IKJ elimination of the rows of one level (Saad, Alg. 10.4); the rows they
// depend on belong to earlier levels and are final.
__global__ void __launch_bounds__(KB) k_ilu_factor(const int *__restrict__ rows, int count, const int *__restrict__ rp,
                                                   const int *__restrict__ col, const int *__restrict__ dg,
                                                   double *__restrict__ a)
{
    const int t = blockIdx.x * KB + threadIdx.x;
    if (t >= count) return;
    const int r = rows[t];
    const int end = rp[r + 1];
    for (int k = rp[r]; k < dg[r]; k++) {
        const int c = col[k];
        const double piv = a[k] / a[dg[c]];
        a[k] = piv;
        int q = k + 1;
        const int cend = rp[c + 1];
        for (int kk = dg[c] + 1; kk < cend; kk++) {
            const int cc = col[kk];
            while (q < end && col[q] < cc) q++;
            if (q < end && col[q] == cc) a[q] -= piv * a[kk];
        }
    }
}

// The same IKJ factorization of a level's rows with FG lanes per row: lane l
// holds the row's entries l, l + FG, ... (FS slots), all pivots' diagonal
// values and the first FG entries of their rows' upper parts are loaded up
// front (the pivot columns are the row's own lower entries, and those rows
// lie in earlier levels), and each update a_rj -= piv * a_cj goes to the lane
// that holds column j, found by comparing the columns passed round the group.
// Every entry receives the same updates in the same order as in
// k_ilu_factor, so the factors are the same to the last bit.  Pivots past
// ILU_FP, or pivot rows with more than FG upper entries, are loaded when
// they are reached.
constexpr int ILU_FG = 16;
constexpr int ILU_FP = 12;
template <int FS>
__device__ __forceinline__ void ilu_factor_pivot(int l, int cp, double pd, int uc, double uv, int u1, int ufirst,
                                                 const int *__restrict__ col, const double *a, int (&mc)[FS],
                                                 double (&mv)[FS], int p)
{
    // the pivot value, as updated by the earlier pivots, from its lane
    double vk = 0.0;
#pragma unroll
    for (int t = 0; t < FS; t++)
        if (t == p / ILU_FG) vk = mv[t];
    vk = __shfl(vk, p % ILU_FG, ILU_FG);
    const double piv = vk / pd;
    if (l == p % ILU_FG) {
#pragma unroll
        for (int t = 0; t < FS; t++)
            if (t == p / ILU_FG) mv[t] = piv;
    }
    // the pivot row's upper entries, ILU_FG at a time
    for (int ub = ufirst;; ub += ILU_FG) {
#pragma unroll
        for (int src = 0; src < ILU_FG; src++) {
            const int c2 = __shfl(uc, src, ILU_FG);
            const double v2 = __shfl(uv, src, ILU_FG);
#pragma unroll
            for (int t = 0; t < FS; t++)
                if (mc[t] == c2) mv[t] -= piv * v2;
        }
        if (ub + ILU_FG >= u1) break;
        const int q = ub + ILU_FG + l;
        uc = q < u1 ? col[q] : -2;
        uv = q < u1 ? a[q] : 0.0;
    }
    (void)cp;
}

template <int FS>
__global__ void __launch_bounds__(KB) k_ilu_factor_rows(const int *__restrict__ rows, int count,
                                                        const int *__restrict__ rp, const int *__restrict__ col,
                                                        const int *__restrict__ dg, double *__restrict__ a)
{
    const int t = blockIdx.x * KB + threadIdx.x;
    const int g = t / ILU_FG, l = t % ILU_FG;
    if (g >= count) return;  // whole groups: KB is a multiple of ILU_FG
    const int r = rows[g], b0 = rp[r], b1 = rp[r + 1], np = dg[r] - b0;
    int mc[FS];
    double mv[FS];
#pragma unroll
    for (int s = 0; s < FS; s++) {
        const int q = b0 + s * ILU_FG + l;
        mc[s] = q < b1 ? col[q] : -1;  // -1: no entry; a pivot row's padding is -2
        mv[s] = q < b1 ? a[q] : 0.0;
    }
    // prefetch: each pivot's column, diagonal value, end of row and the first
    // ILU_FG entries of its upper part (one per lane)
    int pc[ILU_FP], pu1[ILU_FP], puc[ILU_FP];
    double pd[ILU_FP], puv[ILU_FP];
#pragma unroll
    for (int p = 0; p < ILU_FP; p++) {
        pc[p] = 0;
        if (p < np) {
            int cp = 0;
#pragma unroll
            for (int s = 0; s < FS; s++)
                if (s == p / ILU_FG) cp = mc[s];
            pc[p] = __shfl(cp, p % ILU_FG, ILU_FG);
        }
    }
#pragma unroll
    for (int p = 0; p < ILU_FP; p++) {
        pd[p] = 1.0;
        pu1[p] = 0;
        puc[p] = -2;
        puv[p] = 0.0;
        if (p < np) {
            const int dc = dg[pc[p]];
            pu1[p] = rp[pc[p] + 1];
            pd[p] = a[dc];
            const int q = dc + 1 + l;
            puc[p] = q < pu1[p] ? col[q] : -2;
            puv[p] = q < pu1[p] ? a[q] : 0.0;
        }
    }
#pragma unroll
    for (int p = 0; p < ILU_FP; p++)
        if (p < np) ilu_factor_pivot<FS>(l, pc[p], pd[p], puc[p], puv[p], pu1[p], dg[pc[p]] + 1, col, a, mc, mv, p);
    // pivots past the prefetched ones
    for (int p = ILU_FP; p < np; p++) {
        int cp = 0;
#pragma unroll
        for (int s = 0; s < FS; s++)
            if (s == p / ILU_FG) cp = mc[s];
        cp = __shfl(cp, p % ILU_FG, ILU_FG);
        const int dc = dg[cp], u1 = rp[cp + 1], q = dc + 1 + l;
        ilu_factor_pivot<FS>(l, cp, a[dc], q < u1 ? col[q] : -2, q < u1 ? a[q] : 0.0, u1, dc + 1, col, a, mc, mv, p);
    }
#pragma unroll
    for (int s = 0; s < FS; s++) {
        const int q = b0 + s * ILU_FG + l;
        if (q < b1) a[q] = mv[s];
    }
}

// ParILU (Chow & Patel 2015): the ILU(0) equations on the matrix pattern,
//   l_ij u_jj + sum_{k<j} l_ik u_kj = a_ij (i > j),  u_ij + sum_{k<i} l_ik u_kj = a_ij (i <= j),
// solved by synchronous fixed-point sweeps: every entry of a sweep is
// computed from the previous sweep's values (src), so sweeps are
// deterministic (Ginkgo's ParIlu updates in place, asynchronously).  One lane
// per row; u_kj is found in row k right of its diagonal (k < j there).
__global__ void __launch_bounds__(KB) k_parilu_sweep(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                                     const int *__restrict__ dg, const double *__restrict__ a,
                                                     const double *__restrict__ src, double *__restrict__ dst)
{
    const int i = blockIdx.x * KB + threadIdx.x;
    if (i >= n) return;
    const int b = rp[i], e = rp[i + 1];
    for (int q = b; q < e; q++) {
        const int j = col[q], m = min(i, j);
        double acc = a[q];
        for (int q2 = b; q2 < e && col[q2] < m; q2++) {
            const int kk = col[q2];
            for (int t = dg[kk]; t < rp[kk + 1] && col[t] <= j; t++)
                if (col[t] == j) acc -= src[q2] * src[t];
        }
        dst[q] = i > j ? acc / src[dg[j]] : acc;
    }
}

// the sweeps' starting point: l_ij = a_ij / a_jj, u_ij = a_ij
__global__ void __launch_bounds__(KB) k_parilu_init(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                                    const int *__restrict__ dg, const double *__restrict__ a,
                                                    double *__restrict__ lu)
{
    const int i = blockIdx.x * KB + threadIdx.x;
    if (i >= n) return;
    for (int q = rp[i]; q < rp[i + 1]; q++) lu[q] = q < dg[i] ? a[q] / a[dg[col[q]]] : a[q];
}

// sc: flags of the single segment (skip once it has halted) or nullptr
template <typename T>
__global__ void __launch_bounds__(KB) k_ilu_lower(const int *__restrict__ rows, int count, const int *__restrict__ rp,
                                                  const int *__restrict__ col, const int *__restrict__ dg,
                                                  const T *__restrict__ lu, const T *__restrict__ x, T *__restrict__ y,
                                                  const double *sc)
{
    const int t = blockIdx.x * KB + threadIdx.x;
    if (t >= count || (sc && halted(sc))) return;
    const int r = rows[t];
    T s = x[r];
    for (int k = rp[r]; k < dg[r]; k++) s -= lu[k] * y[col[k]];
    y[r] = s;
}

template <typename T>
__global__ void __launch_bounds__(KB) k_ilu_upper(const int *__restrict__ rows, int count, const int *__restrict__ rp,
                                                  const int *__restrict__ col, const int *__restrict__ dg,
                                                  const T *__restrict__ lu, T *__restrict__ y, const double *sc)
{
    const int t = blockIdx.x * KB + threadIdx.x;
    if (t >= count || (sc && halted(sc))) return;
    const int r = rows[t];
    T s = y[r];
    for (int k = dg[r] + 1; k < rp[r + 1]; k++) s -= lu[k] * y[col[k]];
    y[r] = s / lu[dg[r]];
}

// ---------------------------------------------------------------------------
// ILU(0) triangular solves without level launches.  The level-scheduled
// solve above needs one launch per dependency level (about 900 per solve on
// a 64^3 Kuhn mesh at n = 2), and its cost is the GPU-side latency of those
// dependent launches.  Here each triangular solve is one launch:
// * faces are cut into chunks of at most ILU_WG faces (compact boxes of the
//   mesh, else runs of consecutive faces), one work-group each, one lane per
//   face; a lane solves its face's n rows in row order,
//   each row's entries in CSR order: the arithmetic of the row-wise solve,
//   so the results are the same to the last bit;
// * a work-group takes its chunk by ticket (an atomic counter), so chunks
//   start in solve order and every chunk a lane waits on has started before:
//   the earliest unfinished chunk is resident and waits on finished ones;
// * a value is ready when it no longer holds a sentinel (a signalling-NaN
//   payload that arithmetic never produces).  Lanes read values from LDS
//   only.  The chunk's own rows are written there by their lanes; the rows of
//   earlier chunks that the chunk reads (its imports, listed on the host with
//   an LDS slot for every factor entry) are copied in by two more
//   wavefronts, which poll them in the output vector with device-coherent loads
//   (agent-scope atomics).  Producers write the output vector with
//   device-coherent stores.  The value is its own flag, so no fence (which
//   would write back the XCD's L2) is needed;
// * before it waits on anything, a lane loads its face's factors and the
//   LDS slots of its neighbour faces into registers (up to ilu_window<N>()
//   neighbours; the rest are read from memory), so a released wait is
//   followed only by LDS reads and arithmetic: a dependency inside the chunk
//   costs one round of the wavefront, one in another chunk a copy
//   wavefront's poll;
// * each wavefront repeats rounds until its lanes are done.  A lane keeps
//   its place (neighbour, partial sums) when a neighbour is not ready yet
//   and resumes there, polling only that neighbour;
// * every wait is bounded in time: past ILU_WAIT_TICKS of the real-time
//   counter (100 MHz) since the wavefront started, or once another one has
//   given up, the remaining rows (or imports) are written as zero, an error
//   flag is set and the segment is marked DONE, so every launch ends, the
//   queued work of the solve turns into no-ops and the host reports the error.
// The lower solve writes the intermediate vector z and resets the output y
// to the sentinel; the upper solve reads z, writes y and resets z to the
// sentinel for the next apply (each row of z is read only by its own lane).
// Each solve also resets the other's ticket counter.  A matrix that is not
// made of full n x n face blocks, or with a chunk that imports more than
// ILU_IMPORTS values, uses the level launches.
#ifndef MHFEM_ILU_WG
#define MHFEM_ILU_WG 512
#endif
constexpr int ILU_WG = MHFEM_ILU_WG;
constexpr int ILU_NF = 128;                  // the copy wavefronts
constexpr int ILU_TF = 8;                    // imports per copy lane
constexpr int ILU_IMPORTS = ILU_NF * ILU_TF;
constexpr unsigned long long ILU_WAIT_TICKS = 20000000ull;  // 0.2 s at 100 MHz
// neighbour faces whose factors a lane holds (a face of a tetrahedral mesh
// has at most 6 neighbours, 5 of them in the lower triangle at n = 2)
template <int N, bool UPPER>
constexpr int ilu_window()
{
    return N == 1 ? 8 : N == 2 ? (UPPER ? 6 : 5) : 4;
}
// LDS a work-group holds at least: one chunk per CU (160 KiB).  More
// resident chunks only add wavefronts that poll and share the SIMDs with the
// ones on the critical path.  Per step at 64^3 (plate chunks): n = 2 FP64
// 714 ms with two chunks per CU, 634 ms with one; FP32 662 and 570 ms (924 ms
// with the three its registers allowed); n = 1 516 and 413 ms.  With one per
// CU, chunks of 384, 768 or 960 faces were slower than 512 (704, 672, 656 ms).
#ifndef MHFEM_ILU_MIN_LDS
#define MHFEM_ILU_MIN_LDS 83000
#endif
template <typename T, int N, bool UPPER>
constexpr int ilu_lds_values()
{
    constexpr int used = (ILU_WG * N + ILU_IMPORTS) * (int)sizeof(T);
    constexpr int pad = used < MHFEM_ILU_MIN_LDS ? (MHFEM_ILU_MIN_LDS - used) / (int)sizeof(T) : 0;
    return ILU_WG * N + ILU_IMPORTS + pad;
}

template <typename T>
struct SentinelBits;
template <>
struct SentinelBits<double> {
    using U = unsigned long long;
    static constexpr U bits = 0x7FF4A5A5A5A5A5A5ull;
};
template <>
struct SentinelBits<float> {
    using U = unsigned int;
    static constexpr U bits = 0x7FA5A5A5u;
};
template <typename T>
__device__ __forceinline__ bool is_sentinel(T v)
{
    return __builtin_bit_cast(typename SentinelBits<T>::U, v) == SentinelBits<T>::bits;
}
template <typename T>
__device__ __forceinline__ T sentinel_value()
{
    return __builtin_bit_cast(T, SentinelBits<T>::bits);
}
template <typename T>
__device__ __forceinline__ T coherent_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void coherent_store(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T lds_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void lds_store(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct IluSweep {
    int faces, nchunks;
    const int *rp, *dg;
    const int *slot;       // per CSR entry of either triangle: its value's LDS slot
    const int *imp_ptr;    // [0, nchunks] lower solve, [nchunks + 1, 2 nchunks + 1] upper
    const int *imp_rows;   // the imported rows of each chunk, in the order they become ready
    const int *chunk_ptr;  // the faces of chunk c: chunk_face[chunk_ptr[c] .. chunk_ptr[c + 1])
    const int *chunk_face; // [0, faces) lower solve, [faces, 2 faces): in the order of the lanes
    const int *order;      // chunks in start order: [0, nchunks) lower solve, [nchunks, 2 nchunks) upper
    int *ticket;          // [0]: lower solve, [1]: upper solve
    int *err;
    double *sc;  // the scalar slots of every segment: a failed wait marks them DONE
    int nseg;
};

template <typename T>
__global__ void __launch_bounds__(KB) k_ilu_prepare(long n, T *z, int *ticket, int *err)
{
    const T s = sentinel_value<T>();
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) z[i] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ticket[0] = 0;
        ticket[1] = 0;
        *err = 0;
    }
}

__device__ __forceinline__ void ilu_fail(const IluSweep &a)
{
    coherent_store(a.err, 1);
    for (int g = 0; g < a.nseg; g++) a.sc[(long)g * NSLOTS + S_DONE] = 1.0;
}

// lower (UPPER = false: in = x, out = z, reset = y) or upper (UPPER = true:
// in = z, out = y, reset = z) triangular solve.  sc: flags of the single
// segment (the solve is skipped once it has halted; a wait that gives up
// marks it DONE) or nullptr.  ILU_WG solving lanes and ILU_NF copying ones.
template <typename T, int N, bool UPPER>
__global__ void __launch_bounds__(ILU_WG + ILU_NF, 3) k_ilu_sweep(IluSweep a, const T *__restrict__ lu, const T *in, T *out,
                                                               T *reset, const double *sc)
{
    if (sc && halted(sc)) return;
    // the values (and, past them, the padding that holds the work-groups
    // per CU to one; one store keeps it from being taken out)
    __shared__ T ys[ilu_lds_values<T, N, UPPER>()];
    __shared__ int s_ticket;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_ticket = atomicAdd(&a.ticket[UPPER ? 1 : 0], 1);
        // the other solve's counter is free now: reset it for its next launch
        if (s_ticket == 0) a.ticket[UPPER ? 0 : 1] = 0;
    }
    for (int i = tid; i < ILU_WG * N + ILU_IMPORTS; i += ILU_WG + ILU_NF) ys[i] = sentinel_value<T>();
    if (ilu_lds_values<T, N, UPPER>() > ILU_WG * N + ILU_IMPORTS && tid == 0)
        ys[ilu_lds_values<T, N, UPPER>() - 1] = T(0);
    __syncthreads();
    // tickets past the chunk count only follow a failed wait (the counter
    // was not reset): nothing to do
    if (s_ticket >= a.nchunks) return;
    const int chunk = a.order[(UPPER ? a.nchunks : 0) + s_ticket];
    const int c0 = a.chunk_ptr[chunk], nface = a.chunk_ptr[chunk + 1] - c0;
    // a launch that follows a failed wait ends at once
    bool give_up = coherent_load(a.err) != 0;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    int idle = 0;

    if (tid >= ILU_WG) {
        // ---- the copy wavefronts: imports into LDS as they become ready
        T *imp = ys + ILU_WG * N;
        const int j = tid - ILU_WG;
        const int *ip = a.imp_ptr + (UPPER ? a.nchunks + 1 : 0);
        const int i0 = ip[chunk], ni = ip[chunk + 1] - i0;
        int rows[ILU_TF];
        unsigned pend = 0;
#pragma unroll
        for (int t = 0; t < ILU_TF; t++) {
            const int e = j + t * ILU_NF;
            rows[t] = e < ni ? a.imp_rows[i0 + e] : 0;
            if (e < ni) pend |= 1u << t;
        }
        while (__any(pend != 0)) {
            bool progress = false;
            if (pend != 0 && !give_up) {
                // poll the earliest pending import; once it is there, load
                // the lane's other pending ones together
                const int f = __builtin_ctz(pend);
                int row = rows[0];
#pragma unroll
                for (int t = 1; t < ILU_TF; t++)
                    if (t == f) row = rows[t];
                const T v = coherent_load(out + row);
                if (!is_sentinel(v)) {
                    lds_store(&imp[j + f * ILU_NF], v);
                    pend &= pend - 1;
                    progress = true;
                    T w[ILU_TF];
#pragma unroll
                    for (int t = 0; t < ILU_TF; t++) w[t] = (pend >> t & 1u) ? coherent_load(out + rows[t]) : T(0);
#pragma unroll
                    for (int t = 0; t < ILU_TF; t++)
                        if ((pend >> t & 1u) && !is_sentinel(w[t])) {
                            lds_store(&imp[j + t * ILU_NF], w[t]);
                            pend &= ~(1u << t);
                        }
                }
            }
            if (__any(progress)) continue;
            // each poll is a round trip to memory already: a short pause
            __builtin_amdgcn_s_sleep(2);
            if (!give_up && (++idle & 15) == 0 &&
                (__builtin_amdgcn_s_memrealtime() - t0 > ILU_WAIT_TICKS || coherent_load(a.err) != 0))
                give_up = true;
            if (give_up && pend != 0) {
                // release the lanes that wait on these imports
                ilu_fail(a);
#pragma unroll
                for (int t = 0; t < ILU_TF; t++)
                    if (pend >> t & 1u) lds_store(&imp[j + t * ILU_NF], T(0));
                pend = 0;
            }
        }
        return;
    }

    // ---- a solving lane: face E's n rows.  The factors are made of full
    // n x n blocks (checked on the host), so every row of a face lists the
    // same neighbour faces: each neighbour's n values are read once per
    // round and serve all rows.  The chunk's faces are handed to lanes by
    // dependency level, so that a wavefront's lanes become ready together;
    // lane t's face keeps its values at LDS slots t * N + i, and those of a
    // neighbour face from its slot sb on, at sb + j.
    // The factors stay in registers (one work-group per CU leaves room).
    // Rows sum their neighbour terms as the values come in, except the upper
    // solve's rows before the last, which subtract the face's own values
    // first: they are summed at the end, from the neighbour values the lane
    // keeps as it takes them.
    constexpr int K = ilu_window<N, UPPER>();
    bool done = tid >= nface;
    const int E = done ? 0 : a.chunk_face[(UPPER ? a.faces : 0) + c0 + tid];
    const int me = tid * N;
    int nb = 0;          // neighbour faces
    int nbase[N];        // row i's first neighbour entry
    int slotF[K];        // LDS slot of each neighbour face (component 0)
    T lf[N][K][N];       // row i's factor for neighbour b, component j
    T nvs[UPPER ? K : 1][N];  // upper: the neighbour values taken
    T own[N][N];         // the face's own off-diagonal factors (lower: j < i, upper: j > i)
    T z[N], dgv[N];      // right-hand side and (upper) diagonal of each row
#pragma unroll
    for (int i = 0; i < N; i++) {
        nbase[i] = 0;
        z[i] = T(0);
        dgv[i] = T(1);
#pragma unroll
        for (int j = 0; j < N; j++) own[i][j] = T(0);
    }
    if (!done) {
        // a row lists its neighbours and, unless the face's rows are
        // decoupled (a Dirichlet face), the face's own other components
        int d[N], rpv[N + 1];
#pragma unroll
        for (int i = 0; i <= N; i++) rpv[i] = a.rp[E * N + i];
#pragma unroll
        for (int i = 0; i < N; i++) {
            d[i] = a.dg[E * N + i];
            z[i] = in[E * N + i];
        }
        const int nl = d[0] - rpv[0], nu = rpv[N] - d[N - 1] - 1;
        nb = (UPPER ? nu : nl) / N;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const bool has_own = UPPER ? rpv[i + 1] - d[i] - 1 > nu : d[i] - rpv[i] > nl;
            nbase[i] = UPPER ? d[i] + 1 + (has_own ? N - 1 - i : 0) : rpv[i];
            if (UPPER) dgv[i] = lu[d[i]];
#pragma unroll
            for (int j = 0; j < N; j++) {
                if (!UPPER && j < i && has_own) own[i][j] = lu[rpv[i] + nl + j];
                if (UPPER && j > i && has_own) own[i][j] = lu[d[i] + j - i];
            }
        }
    }
#pragma unroll
    for (int b = 0; b < K; b++) {
        const bool has = b < nb;
        slotF[b] = has ? a.slot[nbase[0] + b * N] : 0;
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) lf[i][b][j] = has ? lu[nbase[i] + b * N + j] : T(0);
    }
#pragma unroll
    for (int b = 0; b < (UPPER ? K : 1); b++)
#pragma unroll
        for (int j = 0; j < N; j++) nvs[b][j] = T(0);
    // wait for these loads here, once: the rounds below then have no memory
    // wait on their path (the compiler would otherwise wait in the loop, and
    // so also for the output stores of the previous round)
    __builtin_amdgcn_s_waitcnt(0);
    // partial sums: every row (lower) or the last row (upper: the other rows
    // subtract the face's own values first, so they are summed at the end)
    T s[N];
#pragma unroll
    for (int i = 0; i < N; i++) s[i] = z[i];
    int pos = 0;        // neighbour faces consumed
    bool pend = false;  // waits on the neighbour face at slot ps
    int ps = 0;
    while (__any(!done)) {
        bool progress = false;
        if (!done && pend && !give_up) {
            bool ready = true;
#pragma unroll
            for (int j = 0; j < N; j++) ready = ready && !is_sentinel(lds_load(&ys[ps + j]));
            pend = !ready;
        }
        if (!done && !pend && !give_up) {
            // every neighbour's values at once (past the last neighbour: slot
            // 0, not used), then consumed in CSR order up to the first face
            // not ready, without branches
            T vs[K][N];
#pragma unroll
            for (int b = 0; b < K; b++)
#pragma unroll
                for (int j = 0; j < N; j++) vs[b][j] = lds_load(&ys[slotF[b] + j]);
            const int p0 = pos;
            bool stop = false;
#pragma unroll
            for (int b = 0; b < K; b++) {
                const bool act = !stop && b >= p0 && b < nb;
                bool ready = true;
#pragma unroll
                for (int j = 0; j < N; j++) ready = ready && !is_sentinel(vs[b][j]);
                const bool take = act && ready, blk = act && !ready;
#pragma unroll
                for (int i = 0; i < N; i++) {
                    if (UPPER && i != N - 1) continue;
                    T t = s[i];
#pragma unroll
                    for (int j = 0; j < N; j++) t -= lf[i][b][j] * vs[b][j];
                    s[i] = take ? t : s[i];
                }
                if (UPPER) {
#pragma unroll
                    for (int j = 0; j < N; j++) nvs[UPPER ? b : 0][j] = take ? vs[b][j] : nvs[UPPER ? b : 0][j];
                }
                pos += take ? 1 : 0;
                ps = blk ? slotF[b] : ps;
                stop = stop || blk;
            }
            // neighbours past the register window, from memory
            while (!stop && pos >= K && pos < nb) {
                const int sb = a.slot[nbase[0] + pos * N];
                T lk[N][N];
#pragma unroll
                for (int i = 0; i < N; i++)
#pragma unroll
                    for (int j = 0; j < N; j++) lk[i][j] = lu[nbase[i] + pos * N + j];
                // complete these loads here, so that no later write has to
                // wait for memory
                __builtin_amdgcn_s_waitcnt(0);
                T v[N];
                bool ready = true;
#pragma unroll
                for (int j = 0; j < N; j++) {
                    v[j] = lds_load(&ys[sb + j]);
                    ready = ready && !is_sentinel(v[j]);
                }
                if (!ready) {
                    stop = true;
                    ps = sb;
                } else {
#pragma unroll
                    for (int i = 0; i < N; i++) {
                        if (UPPER && i != N - 1) continue;
#pragma unroll
                        for (int j = 0; j < N; j++) s[i] -= lk[i][j] * v[j];
                    }
                    pos++;
                }
            }
            pend = stop;
            progress = progress || pos != p0;
            if (!stop) {
                // every neighbour is in: the face's rows, in solve order
                T y[N];
                if (!UPPER) {
#pragma unroll
                    for (int i = 0; i < N; i++) {
                        T t = s[i];
#pragma unroll
                        for (int j = 0; j < N; j++)
                            if (j < i) t -= own[i][j] * y[j];
                        y[i] = t;
                    }
                } else {
                    y[N - 1] = s[N - 1] / dgv[N - 1];
#pragma unroll
                    for (int i = N - 2; i >= 0; i--) {
                        T t = z[i];
#pragma unroll
                        for (int j = 0; j < N; j++)
                            if (j > i) t -= own[i][j] * y[j];
                        // the neighbour values, taken in the rounds
#pragma unroll
                        for (int b = 0; b < K; b++) {
                            if (b >= nb) break;
#pragma unroll
                            for (int j = 0; j < N; j++) t -= lf[i][b][j] * nvs[UPPER ? b : 0][j];
                        }
                        for (int b = K; b < nb; b++) {
                            const int sb = a.slot[nbase[0] + b * N];
#pragma unroll
                            for (int j = 0; j < N; j++) t -= lu[nbase[i] + b * N + j] * lds_load(&ys[sb + j]);
                        }
                        y[i] = t / dgv[i];
                    }
                }
#pragma unroll
                for (int i = 0; i < N; i++) {
                    T val = y[i];
                    if (is_sentinel(val)) val = __builtin_nan("");
                    lds_store(&ys[me + i], val);
                    coherent_store(out + E * N + i, val);
                }
                done = true;
                progress = true;
            }
        }
        if (__any(progress)) continue;
        // nothing moved in this wavefront: a short pause leaves the issue
        // slots to the wavefronts that work; give up past the limit
        __builtin_amdgcn_s_sleep(1);
        if (!give_up && (++idle & 31) == 0 &&
            (__builtin_amdgcn_s_memrealtime() - t0 > ILU_WAIT_TICKS || coherent_load(a.err) != 0))
            give_up = true;
        if (give_up && !done) {
            ilu_fail(a);
            // finish the rows with zeros so that nothing waits on them
#pragma unroll
            for (int i = 0; i < N; i++) {
                lds_store(&ys[me + i], T(0));
                coherent_store(out + E * N + i, T(0));
            }
            done = true;
        }
    }
    if (tid < nface) {
#pragma unroll
        for (int i = 0; i < N; i++) reset[E * N + i] = sentinel_value<T>();
    }
}

struct Ilu0 {
    std::vector<int> h_lptr, h_uptr;  // level pointers
    DBuf<int> lrows, urows, diag;     // rows in level order, diagonal positions
    DBuf<double> lu;
    DBuf<float> lu32;
    // single-launch triangular solves: intermediate vector (sentinel between
    // applies), ticket counters and the error flag
    DBuf<double> zmid;
    DBuf<float> zmid32;
    DBuf<int> sweep_state;  // [0], [1]: tickets, [2]: error flag
    DBuf<int> chunk_order;  // start order of the chunks, lower then upper solve
    DBuf<int> slot, imp_ptr, imp_rows;  // LDS slot of every factor entry, each chunk's imports
    DBuf<int> chunk_ptr, chunk_face;    // each chunk's faces in lane order, lower then upper solve
    int nchunks = 0;
    int factor_slots = 0;  // row entries per lane of k_ilu_factor_rows (0: rows too long, k_ilu_factor)
    DBuf<double> a0, swp;  // ParILU: the matrix in CSR order, the other sweep buffer
    // ILUT: factors on their own pattern (rebuilt with the factors)
    bool own = false;
    DBuf<int> frp, fcol;
    std::vector<double> h_a;  // the matrix values, host copy
};

// ---------------------------------------------------------------------------
// ILU-ISAI (the reference's Ginkgo "ILU_ISAI", Solver.hpp:1089-1130): the
// ILU(0) factors L (unit lower) and U are replaced by incomplete sparse
// approximate inverses M_L, M_U on the patterns of L^2 and U^2 (sparsity
// power 2, Solver.hpp:1111), so that the apply y = M_U (M_L x) is two CSR
// products instead of two level-scheduled triangular solves.  Row i of M
// solves (M T)_{i,j} = delta_ij for j in its pattern J (Anzt, Huckle,
// Braeckle, Dongarra 2018), the triangular system T[J,J]^T m = e_i.
//
// One lane per row.  The row's slice of m first holds running sums, then the
// solution: with J sorted, position q is final once every position on its far
// side is (q' > q for L, q' < q for U); row J_q of the factor then adds
// m_q T_{J_q, J_p} to each position p it reaches, found by merging two sorted
// column lists.
template <bool LOWER>
__global__ void __launch_bounds__(KB) k_isai(int n, const int *__restrict__ mrp, const int *__restrict__ mcol,
                                             const int *__restrict__ rp, const int *__restrict__ col,
                                             const int *__restrict__ dg, const double *__restrict__ lu,
                                             double *__restrict__ m)
{
    const int i = blockIdx.x * KB + threadIdx.x;
    if (i >= n) return;
    const int b = mrp[i], e = mrp[i + 1];
    for (int p = b; p < e; p++) m[p] = 0.0;
    for (int s = 0; s < e - b; s++) {
        const int q = LOWER ? e - 1 - s : b + s;
        const int t = mcol[q];
        const double mq = ((t == i ? 1.0 : 0.0) - m[q]) / (LOWER ? 1.0 : lu[dg[t]]);
        m[q] = mq;
        if (LOWER) {  // row t of L left of its diagonal meets positions b..q-1
            int p = b;
            for (int kk = rp[t]; kk < dg[t]; kk++) {
                const int cc = col[kk];
                while (p < q && mcol[p] < cc) p++;
                if (p < q && mcol[p] == cc) m[p] += mq * lu[kk];
            }
        } else {  // row t of U right of its diagonal meets positions q+1..e-1
            int p = q + 1;
            for (int kk = dg[t] + 1; kk < rp[t + 1]; kk++) {
                const int cc = col[kk];
                while (p < e && mcol[p] < cc) p++;
                if (p < e && mcol[p] == cc) m[p] += mq * lu[kk];
            }
        }
    }
}

// y = M x, CSR with L lanes per row (L is measured once per matrix: the best
// choice depends on how the row lengths spread); sc: flags of the single
// segment (skip once it has halted) or nullptr
template <int L>
__global__ void __launch_bounds__(KB) k_csr_mv(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                               const double *__restrict__ val, const double *__restrict__ x,
                                               double *__restrict__ y, const double *sc)
{
    if (sc && halted(sc)) return;
    const long t = (long)blockIdx.x * KB + threadIdx.x;
    const int row = (int)(t / L), lane = (int)(t % L);
    double s = 0.0;
    if (row < n)
        for (int q = rp[row] + lane; q < rp[row + 1]; q += L) s += val[q] * x[col[q]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (row < n && lane == 0) y[row] = s;
}

struct IsaiFactor {
    DBuf<int> rp, col;  // pattern (that of L^2 or U^2), CSR
    DBuf<double> val;
    long nnz = 0;
    int lanes = 8;  // lanes per row of the product, measured at the first build
};

struct Isai {
    IsaiFactor lo, up;  // M_L, M_U
    DBuf<double> t;     // M_L x
    // reuse rule (Solver.hpp:1017, 1158-1160)
    long long last_iters = 0, updated_iters = 0;
    int rebuilds = 0;
    bool rebuilt = false;
};

// ---------------------------------------------------------------------------
struct Krylov {
    int n = 0;
    DBuf<double> r, r_ast, p, Ap, s, As, tmp, tmp2, dinv;
    // segmented decomposition
    int nseg = 1, rows_per_block = 2048;
    DBuf<int> blk_seg, blk_row, seg_blk;
    std::vector<int> h_seg_blk;  // host copy of seg_blk
    DBuf<int> active;            // nblk: the blocks of the running segments
    int nblk = 0;
    DBuf<double> partials;
    DBuf<double> sc;
    double *h_sc = nullptr;  // pinned, nseg * NSLOTS
    // flags of a batch of iterations read while the next batch runs
    double *h_sc_ahead[2] = {nullptr, nullptr};  // pinned, nseg * NSLOTS each
    hipEvent_t ev_ahead[2] = {nullptr, nullptr};
    Ilu0 ilu;
    Isai isai;
    Amg *amg = nullptr;
    // FP32 inner solver
    DBuf<float> val32, dinv32, b32, x32, r32, r_ast32, p32, Ap32, s32, As32, tmp32, rhs32;
    // its matrix in a face order that groups faces with the same neighbour
    // offsets (see build_fp32_order): perm[p] = natural face, iperm inverse
    bool permuted = false;
    // the same order for the FP64 iteration: permuted FP64 matrix copy, the
    // right-hand side and iterate of the solve in that order
    bool permuted64 = false;
    DBuf<double> val64p, bp, xp;
    DBuf<int> perm, iperm, slice_width32, col32;
    DBuf<int64_t> slice_ptr32, map32;
    int64_t blocks32 = 0, diag_base32 = 0;
    // GMRES
    DBuf<double> V;
    int V_m = 0;
    // vectors of TFQMR, BiCGStab(l) and IDR(s) beyond the BiCGStab set
    // (allocated on first use)
    DBuf<double> ext;
    // mixed precision: solves left that go straight to FP64 after a fallback
    int fp64_cooldown = 0;
    // captured step graphs, and the scheme arguments they were captured with
    hipGraphExec_t g_first = nullptr, g_more = nullptr, g_post = nullptr;
    double g_tau = 0;
    int g_iters = 0;
    SchemeView g_view{};
    hipStream_t g_stream = nullptr;
};

static inline unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + KB - 1) / KB, MAXG)); }

// work-group tables for `rpb` rows per work-group, segment aligned
static void build_blocks(Ctx &c, Krylov &k, int rpb)
{
    // whole faces per work-group: the matrix kernels run one lane per face
    rpb = std::max(c.n, rpb / c.n * c.n);
    std::vector<int> bseg, brow, sblk(1, 0);
    for (int s = 0; s < k.nseg; s++) {
        const int r0 = c.h_seg_rows[s], r1 = c.h_seg_rows[s + 1];
        for (int r = r0; r < r1; r += rpb) {
            bseg.push_back(s);
            brow.push_back(r);
        }
        if (r1 == r0) {  // empty segment: one empty work-group keeps the tables uniform
            bseg.push_back(s);
            brow.push_back(r0);
        }
        sblk.push_back((int)bseg.size());
    }
    brow.push_back(c.h_seg_rows[k.nseg]);
    // brow[b + 1] is the end of work-group b: the next work-group starts where
    // this one ends, and every segment's first work-group starts at its first row
    k.nblk = (int)bseg.size();
    k.rows_per_block = rpb;
    k.blk_seg.alloc(k.nblk);
    k.blk_row.alloc(k.nblk + 1);
    k.seg_blk.alloc(k.nseg + 1);
    MHFEM_HIP(hipMemcpy(k.blk_seg.p, bseg.data(), k.nblk * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.blk_row.p, brow.data(), (k.nblk + 1) * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.seg_blk.p, sblk.data(), (k.nseg + 1) * sizeof(int), hipMemcpyHostToDevice));
    k.partials.alloc((size_t)4 * k.nblk);
    k.active.alloc(k.nblk);
    k.h_seg_blk = sblk;
}

// Face order of the FP32 iteration's matrix copy.  The natural
// (first-appearance) numbering interleaves the face types of a cell, so the
// 64 faces of a slice gather x from 64 unrelated places.  Faces whose stored
// blocks sit at the same offsets F - E (faces of one type away from the
// boundary) are taken together inside windows of the natural order: the k-th
// column faces of consecutive faces of a group are then mostly consecutive
// faces of one other group, and a wavefront's gathers fall into few cache
// lines.  The FP32 vectors live in this order;
// the FP64 residual and correction kernels translate at the boundary.
struct OffsetsHash {
    size_t operator()(const std::vector<int> &v) const
    {
        uint64_t h = 1469598103934665603ull;
        for (int x : v) h = (h ^ (uint32_t)x) * 1099511628211ull;
        return (size_t)h;
    }
};

static void build_fp32_order(Ctx &c, Krylov &k)
{
    const int F = c.F, n = c.n, nn = n * n;
    std::unordered_map<std::vector<int>, int, OffsetsHash> groups;
    std::vector<int> gid(F);
    std::vector<int> key;
    for (int E = 0; E < F; E++) {
        key.clear();
        for (int64_t q = c.h_bptr[E]; q < c.h_bptr[E + 1]; q++) key.push_back(c.h_bface[q] - E);
        auto it = groups.find(key);
        if (it == groups.end()) it = groups.emplace(key, (int)groups.size()).first;
        gid[E] = it->second;
    }
    // within each window of CHUNK consecutive faces, a stable sort by group:
    // the windows keep the natural order's locality (each XCD sweeps a
    // contiguous range of faces and keeps its part of x in its L2), the
    // groups inside a window give the coalesced runs
    constexpr int CHUNK = 4096;
    std::vector<int> perm(F), iperm(F), order;
    for (int c0 = 0; c0 < F; c0 += CHUNK) {
        const int c1 = std::min(F, c0 + CHUNK);
        order.resize(c1 - c0);
        for (int E = c0; E < c1; E++) order[E - c0] = E;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return gid[a] < gid[b]; });
        std::copy(order.begin(), order.end(), perm.begin() + c0);
    }
    for (int p = 0; p < F; p++) iperm[perm[p]] = p;
    // block-SELL layout of the permuted matrix
    const int nslices = (F + 63) / 64;
    std::vector<int> w32(nslices, 0);
    std::vector<int64_t> sp32(nslices + 1, 0);
    for (int s = 0; s < nslices; s++) {
        int w = 0;
        for (int p = s * 64; p < std::min(F, s * 64 + 64); p++)
            w = std::max<int>(w, (int)(c.h_bptr[perm[p] + 1] - c.h_bptr[perm[p]]));
        w32[s] = w;
        sp32[s + 1] = sp32[s] + (int64_t)w * 64;
    }
    const int64_t blocks = sp32[nslices];
    std::vector<int> col(std::max<int64_t>(blocks, 1));
    std::vector<int64_t> map(std::max<int64_t>(blocks, 1), -1);
    for (int s = 0; s < nslices; s++)
        for (int lane = 0; lane < 64; lane++) {
            const int p = s * 64 + lane;
            const int E = p < F ? perm[p] : -1;
            const int64_t cnt = E >= 0 ? c.h_bptr[E + 1] - c.h_bptr[E] : 0;
            for (int kk = 0; kk < w32[s]; kk++) {
                const int64_t b = sell_block(sp32.data(), p, kk);
                if (kk < cnt) {
                    col[b] = iperm[c.h_bface[c.h_bptr[E] + kk]];
                    map[b] = sell_block(c.h_slice_ptr.data(), E, kk);
                } else {
                    col[b] = p < F ? p : 0;
                    map[b] = -1;
                }
            }
        }
    for (int64_t b = 0; b < blocks; b++)
        if (col[b] < 0 || col[b] >= F || map[b] >= c.sell_blocks)
            throw Error{MHFEM_E_STATE, "inconsistent FP32 matrix order"};
    k.blocks32 = blocks;
    k.diag_base32 = blocks * nn;
    k.perm.alloc(F);
    k.iperm.alloc(F);
    k.slice_width32.alloc(nslices);
    k.slice_ptr32.alloc(nslices + 1);
    k.col32.alloc(col.size());
    k.map32.alloc(map.size());
    MHFEM_HIP(hipMemcpy(k.perm.p, perm.data(), F * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.iperm.p, iperm.data(), F * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.slice_width32.p, w32.data(), nslices * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.slice_ptr32.p, sp32.data(), (nslices + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.col32.p, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(k.map32.p, map.data(), map.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    k.permuted = true;
}

static void isai_mv(Ctx &c, const IsaiFactor &F, const double *x, double *y, const double *flag, int lanes = 0)
{
    const int n = c.rows;
    auto go = [&](auto LC) {
        constexpr int L = decltype(LC)::value;
        hipLaunchKernelGGL((k_csr_mv<L>), (unsigned)(((long)n * L + KB - 1) / KB), KB, 0, c.stream, n, F.rp.p,
                           F.col.p, F.val.p, x, y, flag);
    };
    switch (lanes ? lanes : F.lanes) {
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    default: go(std::integral_constant<int, 32>{}); break;
    }
    MHFEM_HIP(hipGetLastError());
}

// the fastest lanes per row for this matrix: each candidate timed over three
// launches after one warm-up launch (x, y: any vectors of the system's length)
static int isai_tune_lanes(Ctx &c, const IsaiFactor &F, const double *x, double *y)
{
    hipEvent_t e0, e1;
    MHFEM_HIP(hipEventCreate(&e0));
    MHFEM_HIP(hipEventCreate(&e1));
    int best = 8;
    float best_ms = 1e30f;
    for (int lanes : {2, 4, 8, 16, 32}) {
        isai_mv(c, F, x, y, nullptr, lanes);
        MHFEM_HIP(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 3; r++) isai_mv(c, F, x, y, nullptr, lanes);
        MHFEM_HIP(hipEventRecord(e1, c.stream));
        MHFEM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MHFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) best_ms = ms, best = lanes;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// level tables of a factor pattern (sorted columns, diagonal present): the
// rows of the lower and of the upper triangular solve bucketed by dependency
// depth, and the diagonal positions; returns the latter
// Chunks of the single-launch triangular solves from the faces' positions
// (the lowest corner of a face's first cell: all cells of a subdivided cube
// share it, so a cube's faces fall in one box).  The boxes are plates one
// cell layer thick across the direction in which neighbouring faces lie
// farthest apart in the face order (the outer loop of a lexicographic
// numbering), sized for about 7/8 of ILU_WG faces (a fuller box is cut in
// pieces of at most ILU_WG in face order).  A plate keeps the dependencies
// along the inner directions inside one chunk, in LDS; on the 64^3 Kuhn mesh
// such plates ran the solves 20 % faster than runs of consecutive faces (a
// piece of a row of cubes), cubic boxes 15 % slower, plates across another
// direction 50 % slower.  The chunk dependency graph must have no cycle
// (every chunk's imports come from chunks started before it); returns false,
// and the caller takes runs of consecutive faces, when it has one, when no
// positions are known or when the faces do not lie in layers.
static bool ilu_box_chunks(const Ctx &c, const std::vector<int> &dg, std::vector<int> &chunk_of, int &nch)
{
    const int F = c.F, n = c.n;
    if (c.h_face_pos.size() != 3 * (size_t)F || F < 2 * ILU_WG) return false;
    const float *pos = c.h_face_pos.data();
    float lo[3], hi[3];
    for (int d = 0; d < 3; d++) lo[d] = hi[d] = pos[d];
    for (int E = 0; E < F; E++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], pos[3 * (size_t)E + d]);
            hi[d] = std::max(hi[d], pos[3 * (size_t)E + d]);
        }
    // per direction: the face-order gaps and the smallest distance between
    // a face and a lower neighbour displaced along that direction only
    std::vector<int> gaps[3];
    double spacing[3] = {1e300, 1e300, 1e300};
    for (int E = 0; E < F; E++)
        for (int q = c.h_rowptr[E * n]; q < dg[E * n]; q++) {
            const int G = c.h_col[q] / n;
            int along = -1, moved = 0;
            for (int d = 0; d < 3; d++) {
                const double dp = std::fabs((double)pos[3 * (size_t)E + d] - pos[3 * (size_t)G + d]);
                if (dp > 1e-6 * (hi[d] - lo[d] + 1e-30)) {
                    along = d;
                    moved++;
                    spacing[d] = std::min(spacing[d], dp);
                }
            }
            if (moved == 1) gaps[along].push_back(E - G);
        }
    int outer = -1;
    double best = -1.0;
    for (int d = 0; d < 3; d++) {
        if (gaps[d].empty() || !(hi[d] > lo[d])) continue;
        auto mid = gaps[d].begin() + gaps[d].size() / 2;
        std::nth_element(gaps[d].begin(), mid, gaps[d].end());
        if (*mid > best) {
            best = *mid;
            outer = d;
        }
    }
    if (outer < 0) return false;
    const double layers = std::floor((hi[outer] - lo[outer]) / spacing[outer] + 0.5) + 1.0;
    double area = 1.0;
    int inner = 0;
    for (int d = 0; d < 3; d++)
        if (d != outer && hi[d] > lo[d]) {
            area *= hi[d] - lo[d];
            inner++;
        }
    if (inner == 0) return false;
    // side of the plate in the inner directions
    const double h = std::pow(area * layers * (0.875 * ILU_WG) / F, 1.0 / inner);
    int g[3];
    for (int d = 0; d < 3; d++) {
        if (!(hi[d] > lo[d])) g[d] = 1;
        else if (d == outer) g[d] = std::max(1, (int)std::ceil((hi[d] - lo[d]) / (0.8 * spacing[d])));
        else g[d] = std::max(1, (int)std::ceil((hi[d] - lo[d]) / h));
    }
    if ((double)g[0] * g[1] * g[2] > 4.0 * F) return false;
    std::vector<int> box(F);
    std::vector<int> cnt((size_t)g[0] * g[1] * g[2], 0);
    for (int E = 0; E < F; E++) {
        int id = 0;
        for (int d = 2; d >= 0; d--) {
            const double t = g[d] > 1 ? (c.h_face_pos[3 * (size_t)E + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
            const int q = std::min(g[d] - 1, std::max(0, (int)(t * g[d])));
            id = id * g[d] + q;
        }
        box[E] = id;
        cnt[id]++;
    }
    // chunk ids: boxes in grid order, a box cut in pieces of at most ILU_WG
    std::vector<int> first(cnt.size() + 1, 0), seen(cnt.size(), 0);
    for (size_t b = 0; b < cnt.size(); b++) first[b + 1] = first[b] + (cnt[b] + ILU_WG - 1) / ILU_WG;
    nch = first.back();
    chunk_of.resize(F);
    for (int E = 0; E < F; E++) chunk_of[E] = first[box[E]] + seen[box[E]]++ / ILU_WG;
    // the chunk graph of the lower solve (the upper one is its reverse) must
    // be acyclic: Kahn's algorithm over its distinct edges
    std::vector<std::vector<int>> succ(nch);
    for (int E = 0; E < F; E++) {
        const int r = E * n;
        for (int q = c.h_rowptr[r]; q < dg[r]; q++) {
            const int from = chunk_of[c.h_col[q] / n];
            if (from != chunk_of[E]) succ[from].push_back(chunk_of[E]);
        }
    }
    std::vector<int> indeg(nch, 0), queue;
    for (auto &v : succ) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        for (int t : v) indeg[t]++;
    }
    for (int ch = 0; ch < nch; ch++)
        if (indeg[ch] == 0) queue.push_back(ch);
    for (size_t qi = 0; qi < queue.size(); qi++)
        for (int t : succ[queue[qi]])
            if (--indeg[t] == 0) queue.push_back(t);
    return (int)queue.size() == nch;
}

// Tables of the single-launch triangular solves: the chunks (box chunks,
// else runs of ILU_WG consecutive faces), each chunk's faces in the order of
// its lanes (by level in the face dependency graph of each triangle), the
// faces of other chunks each chunk reads (its imports, all n rows of each:
// ascending for the lower solve, descending for the upper one, the order in
// which they are written), for every factor entry the LDS slot its value is
// read from (the chunk's own faces by lane, then its imports), and the start
// order of the chunks (by depth in the chunk dependency graph).  False, and
// nothing uploaded, when the pattern is not made of full n x n face blocks
// (the solve reads a neighbour face's n values together) or a chunk imports
// more than ILU_IMPORTS values.
static bool ilu_sweep_tables(const Ctx &c, Ilu0 &I, const std::vector<int> &dg)
{
    const int n = c.n, F = c.F;
    const int *rp = c.h_rowptr.data(), *col = c.h_col.data();
    // every row of face E: lower part [neighbours][own j < i][diagonal],
    // upper part [diagonal][own j > i][neighbours], the same neighbour list
    // (full blocks of ascending faces) in every row, and the face's own
    // components either all there or none (rows decoupled: a Dirichlet face)
    for (int E = 0; E < F; E++) {
        const int r0 = E * n, r1 = E * n + n - 1;
        const int nl = dg[r0] - rp[r0], nu = rp[r1 + 1] - dg[r1] - 1;
        if (nl % n || nu % n) return false;
        for (int part = 0; part < 2; part++) {
            const int *list = col + (part == 0 ? rp[r0] : dg[r1] + 1), len = part == 0 ? nl : nu;
            for (int q = 0; q < len; q++)
                if (list[q] % n != q % n || list[q] / n != list[q - q % n] / n || list[q] / n == E) return false;
        }
        const bool own = dg[r1] - rp[r1] > nl;
        for (int i = 0; i < n; i++) {
            const int r = E * n + i, ol = own ? i : 0, ou = own ? n - 1 - i : 0;
            if (dg[r] - rp[r] != nl + ol || rp[r + 1] - dg[r] - 1 != ou + nu) return false;
            for (int q = 0; q < nl; q++)
                if (col[rp[r] + q] != col[rp[r0] + q]) return false;
            for (int j = 0; j < ol; j++)
                if (col[rp[r] + nl + j] != E * n + j) return false;
            for (int j = 1; j <= ou; j++)
                if (col[dg[r] + j] != E * n + i + j) return false;
            for (int q = 0; q < nu; q++)
                if (col[dg[r] + 1 + ou + q] != col[dg[r1] + 1 + q]) return false;
        }
    }
    // chunks
    std::vector<int> chunk_of;
    int nch = 0;
    if (!ilu_box_chunks(c, dg, chunk_of, nch)) {
        nch = (F + ILU_WG - 1) / ILU_WG;
        chunk_of.resize(F);
        for (int E = 0; E < F; E++) chunk_of[E] = E / ILU_WG;
    }
    std::vector<int> cptr(nch + 1, 0);
    for (int E = 0; E < F; E++) cptr[chunk_of[E] + 1]++;
    for (int ch = 0; ch < nch; ch++) cptr[ch + 1] += cptr[ch];
    std::vector<int> members(F), fill(cptr.begin(), cptr.end() - 1);
    for (int E = 0; E < F; E++) members[fill[chunk_of[E]]++] = E;  // ascending within a chunk
    // per triangle: face levels, the lanes' order, the lane of each face
    std::vector<int> lev(F), cface(2 * (size_t)F), lane(2 * (size_t)F);
    for (int pass = 0; pass < 2; pass++) {
        for (int t = 0; t < F; t++) {
            const int E = pass == 0 ? t : F - 1 - t;
            int l = 0;
            for (int r = E * n; r < (E + 1) * n; r++) {
                const int q0 = pass == 0 ? rp[r] : dg[r] + 1, q1 = pass == 0 ? dg[r] : rp[r + 1];
                for (int q = q0; q < q1; q++) {
                    const int G = col[q] / n;
                    if (G != E) l = std::max(l, lev[G] + 1);
                }
            }
            lev[E] = l;
        }
        int *L = cface.data() + (size_t)pass * F, *ln = lane.data() + (size_t)pass * F;
        for (int ch = 0; ch < nch; ch++) {
            std::copy(members.begin() + cptr[ch], members.begin() + cptr[ch + 1], L + cptr[ch]);
            std::stable_sort(L + cptr[ch], L + cptr[ch + 1], [&](int x, int y) { return lev[x] < lev[y]; });
            for (int t = cptr[ch]; t < cptr[ch + 1]; t++) ln[L[t]] = t - cptr[ch];
        }
    }
    // imports and slots
    std::vector<int> slot(c.nnz, 0), iptr(2 * ((size_t)nch + 1)), irows, mark(F, -1), where(F, 0);
    std::vector<int> list;
    for (int pass = 0; pass < 2; pass++) {
        int *P = iptr.data() + (size_t)pass * (nch + 1);
        const int *ln = lane.data() + (size_t)pass * F;
        for (int ch = 0; ch < nch; ch++) {
            const int key = pass * nch + ch;
            P[ch] = (int)irows.size();
            list.clear();
            for (int t = cptr[ch]; t < cptr[ch + 1]; t++)
                for (int r = members[t] * n; r < (members[t] + 1) * n; r++) {
                    const int q0 = pass == 0 ? rp[r] : dg[r] + 1, q1 = pass == 0 ? dg[r] : rp[r + 1];
                    for (int q = q0; q < q1; q++) {
                        const int G = col[q] / n;
                        if (chunk_of[G] == ch) {
                            slot[q] = ln[G] * n + col[q] % n;
                        } else if (mark[G] != key) {
                            mark[G] = key;
                            list.push_back(G);
                        }
                    }
                }
            if (list.size() * n > (size_t)ILU_IMPORTS) return false;
            if (pass == 0) std::sort(list.begin(), list.end());
            else std::sort(list.begin(), list.end(), std::greater<int>());
            for (size_t e = 0; e < list.size(); e++) where[list[e]] = (int)e;
            for (int t = cptr[ch]; t < cptr[ch + 1]; t++)
                for (int r = members[t] * n; r < (members[t] + 1) * n; r++) {
                    const int q0 = pass == 0 ? rp[r] : dg[r] + 1, q1 = pass == 0 ? dg[r] : rp[r + 1];
                    for (int q = q0; q < q1; q++) {
                        const int G = col[q] / n;
                        if (chunk_of[G] != ch) slot[q] = ILU_WG * n + where[G] * n + col[q] % n;
                    }
                }
            for (const int G : list)
                for (int j = 0; j < n; j++) irows.push_back(G * n + j);
        }
        P[nch] = (int)irows.size();
    }
    // start order: by depth in the chunk dependency graph, then by index.
    // Each chunk's depth follows from those of the chunks it imports from,
    // which lie earlier in a topological order of the (acyclic) graph
    std::vector<int> ord(2 * (size_t)nch), depth(nch), topo;
    {
        std::vector<std::vector<int>> pred(nch);
        for (int E = 0; E < F; E++)
            for (int q = rp[E * n]; q < dg[E * n]; q++) {
                const int from = chunk_of[col[q] / n];
                if (from != chunk_of[E]) pred[chunk_of[E]].push_back(from);
            }
        std::vector<int> indeg(nch, 0);
        std::vector<std::vector<int>> succ(nch);
        for (int ch = 0; ch < nch; ch++) {
            std::sort(pred[ch].begin(), pred[ch].end());
            pred[ch].erase(std::unique(pred[ch].begin(), pred[ch].end()), pred[ch].end());
            indeg[ch] = (int)pred[ch].size();
            for (int f : pred[ch]) succ[f].push_back(ch);
        }
        for (int ch = 0; ch < nch; ch++)
            if (indeg[ch] == 0) topo.push_back(ch);
        for (size_t qi = 0; qi < topo.size(); qi++)
            for (int t : succ[topo[qi]])
                if (--indeg[t] == 0) topo.push_back(t);
        if ((int)topo.size() != nch) return false;  // cannot happen for runs of consecutive faces
        for (int pass = 0; pass < 2; pass++) {
            std::fill(depth.begin(), depth.end(), 0);
            for (int t = 0; t < nch; t++) {
                const int ch = pass == 0 ? topo[t] : topo[nch - 1 - t];
                for (int f : pass == 0 ? pred[ch] : succ[ch]) depth[ch] = std::max(depth[ch], depth[f] + 1);
            }
            int *o = ord.data() + (size_t)pass * nch;
            std::iota(o, o + nch, 0);
            if (pass == 1) std::reverse(o, o + nch);
            std::stable_sort(o, o + nch, [&](int x, int y) { return depth[x] < depth[y]; });
        }
    }
    I.nchunks = nch;
    I.chunk_ptr.alloc(cptr.size());
    I.chunk_face.alloc(cface.size());
    I.chunk_order.alloc(ord.size());
    I.slot.alloc(slot.size());
    I.imp_ptr.alloc(iptr.size());
    I.imp_rows.alloc(std::max<size_t>(irows.size(), 1));
    MHFEM_HIP(hipMemcpy(I.chunk_ptr.p, cptr.data(), cptr.size() * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(I.chunk_face.p, cface.data(), cface.size() * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(I.chunk_order.p, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(I.slot.p, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(I.imp_ptr.p, iptr.data(), iptr.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!irows.empty())
        MHFEM_HIP(hipMemcpy(I.imp_rows.p, irows.data(), irows.size() * sizeof(int), hipMemcpyHostToDevice));
    return true;
}

static std::vector<int> build_levels(Ilu0 &I, int n, const int *rp, const int *col)
{
    std::vector<int> dg(n), lev(n), ulev(n);
    for (int r = 0; r < n; r++)
        for (int q = rp[r]; q < rp[r + 1]; q++)
            if (col[q] == r) dg[r] = q;
    int nl = 0, nu = 0;
    for (int r = 0; r < n; r++) {
        int l = 0;
        for (int q = rp[r]; q < dg[r]; q++) l = std::max(l, lev[col[q]] + 1);
        lev[r] = l;
        nl = std::max(nl, l + 1);
    }
    for (int r = n - 1; r >= 0; r--) {
        int l = 0;
        for (int q = dg[r] + 1; q < rp[r + 1]; q++) l = std::max(l, ulev[col[q]] + 1);
        ulev[r] = l;
        nu = std::max(nu, l + 1);
    }
    auto bucket = [n](const std::vector<int> &lv, int nlev, std::vector<int> &ptr, std::vector<int> &rows) {
        ptr.assign(nlev + 1, 0);
        for (int r = 0; r < n; r++) ptr[lv[r] + 1]++;
        for (int l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
        rows.resize(n);
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int r = 0; r < n; r++) rows[pos[lv[r]]++] = r;
    };
    std::vector<int> lrows, urows;
    bucket(lev, nl, I.h_lptr, lrows);
    bucket(ulev, nu, I.h_uptr, urows);
    if (!I.lrows.p) {
        I.lrows.alloc(n);
        I.urows.alloc(n);
        I.diag.alloc(n);
    }
    MHFEM_HIP(hipMemcpy(I.lrows.p, lrows.data(), n * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(I.urows.p, urows.data(), n * sizeof(int), hipMemcpyHostToDevice));
    MHFEM_HIP(hipMemcpy(I.diag.p, dg.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return dg;
}

// ILUT(p, tau) (Saad, Iterative Methods for Sparse Linear Systems, Alg. 10.6;
// TNL's "ilut" preconditioner), on the host: each row of A, in a dense work
// row, is eliminated against the finished rows of U in increasing column
// order (fill-in included); multipliers below tau_i = tau ||a_i||_2 are
// dropped; of the rest, entries below tau_i go and only the p largest of the
// L part and of the U part stay (ties to the smaller column), the diagonal
// always (a zero pivot becomes tau_i, or 1).  Output: one CSR, sorted
// columns, unit-lower L left of U's diagonal.
static void ilut_host(int n, const int *rp, const int *col, const double *val, int p, double tau,
                      std::vector<int> &frp, std::vector<int> &fcol, std::vector<double> &fval)
{
    std::vector<double> w(n, 0.0);
    std::vector<int> mark(n, -1), seen(n, -1), fdg(n), nz;
    std::vector<std::pair<int, double>> lo, up;
    frp.assign(1, 0);
    fcol.clear();
    fval.clear();
    fcol.reserve((size_t)rp[n] * 2);
    fval.reserve((size_t)rp[n] * 2);
    auto keep = [p](std::vector<std::pair<int, double>> &v) {
        std::sort(v.begin(), v.end(), [](const auto &a, const auto &b) {
            const double x = std::fabs(a.second), y = std::fabs(b.second);
            return x != y ? x > y : a.first < b.first;
        });
        if ((int)v.size() > p) v.resize(p);
        std::sort(v.begin(), v.end());
    };
    for (int i = 0; i < n; i++) {
        nz.clear();
        double nrm = 0.0;
        auto touch = [&](int j) {
            if (mark[j] != i) {
                mark[j] = i;
                w[j] = 0.0;
                nz.push_back(j);
            }
        };
        for (int q = rp[i]; q < rp[i + 1]; q++) {
            touch(col[q]);
            w[col[q]] = val[q];
            nrm += val[q] * val[q];
        }
        touch(i);
        const double ti = tau * std::sqrt(nrm);
        for (;;) {
            int kk = -1;
            for (int j : nz)
                if (j < i && seen[j] != i && (kk < 0 || j < kk)) kk = j;
            if (kk < 0) break;
            seen[kk] = i;
            if (w[kk] == 0.0) continue;
            w[kk] /= fval[fdg[kk]];
            if (std::fabs(w[kk]) < ti) {
                w[kk] = 0.0;
                continue;
            }
            for (int q = fdg[kk] + 1; q < frp[kk + 1]; q++) {
                touch(fcol[q]);
                w[fcol[q]] -= w[kk] * fval[q];
            }
        }
        lo.clear();
        up.clear();
        for (int j : nz)
            if (j != i && w[j] != 0.0 && std::fabs(w[j]) >= ti) (j < i ? lo : up).emplace_back(j, w[j]);
        keep(lo);
        keep(up);
        for (const auto &e : lo) fcol.push_back(e.first), fval.push_back(e.second);
        fdg[i] = (int)fcol.size();
        fcol.push_back(i);
        fval.push_back(w[i] != 0.0 ? w[i] : (ti > 0 ? ti : 1.0));
        for (const auto &e : up) fcol.push_back(e.first), fval.push_back(e.second);
        if (fcol.size() > (size_t)INT32_MAX) throw Error{MHFEM_E_ARGUMENT, "ILUT factors exceed 2^31 entries"};
        frp.push_back((int)fcol.size());
    }
}

// patterns of L^2 and U^2 from the factors' pattern (that of the matrix):
// row i of L^2 is the union of the rows t of L that row i of L reaches
static void isai_patterns(Ctx &c, Krylov &k, const std::vector<int> &dg)
{
    const int n = k.n;
    const std::vector<int> &rp = c.h_rowptr, &col = c.h_col;
    auto build = [&](bool lower, IsaiFactor &F) {
        std::vector<int> prp(n + 1, 0), pcol, mark(n, -1), row;
        pcol.reserve((size_t)rp[n] * 2);
        for (int i = 0; i < n; i++) {
            row.clear();
            for (int a = lower ? rp[i] : dg[i]; a < (lower ? dg[i] + 1 : rp[i + 1]); a++) {
                const int t = col[a];
                for (int q = lower ? rp[t] : dg[t]; q < (lower ? dg[t] + 1 : rp[t + 1]); q++)
                    if (mark[col[q]] != i) {
                        mark[col[q]] = i;
                        row.push_back(col[q]);
                    }
            }
            std::sort(row.begin(), row.end());
            pcol.insert(pcol.end(), row.begin(), row.end());
            if (pcol.size() > (size_t)INT32_MAX) throw Error{MHFEM_E_ARGUMENT, "ILU-ISAI pattern exceeds 2^31 entries"};
            prp[i + 1] = (int)pcol.size();
        }
        F.nnz = (long)pcol.size();
        F.rp.alloc(n + 1);
        F.col.alloc(F.nnz);
        F.val.alloc(F.nnz);
        MHFEM_HIP(hipMemcpy(F.rp.p, prp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(F.col.p, pcol.data(), F.nnz * sizeof(int), hipMemcpyHostToDevice));
    };
    Isai &S = k.isai;
    build(true, S.lo);
    build(false, S.up);
    S.t.alloc(n);
}

Krylov *krylov_create(Ctx &c)
{
    Krylov *k = new Krylov();
    const int n = c.rows;
    k->n = n;
    k->nseg = (int)c.h_seg_rows.size() - 1;
    for (DBuf<double> *b : {&k->r, &k->r_ast, &k->p, &k->Ap, &k->s, &k->As, &k->tmp, &k->tmp2, &k->dinv})
        b->alloc(n);
    k->sc.alloc((size_t)k->nseg * NSLOTS);
    MHFEM_HIP(hipHostMalloc((void **)&k->h_sc, (size_t)k->nseg * NSLOTS * sizeof(double), hipHostMallocDefault));
    for (int q = 0; q < 2; q++) {
        MHFEM_HIP(hipHostMalloc((void **)&k->h_sc_ahead[q], (size_t)k->nseg * NSLOTS * sizeof(double),
                                hipHostMallocDefault));
        MHFEM_HIP(hipEventCreateWithFlags(&k->ev_ahead[q], hipEventDisableTiming));
    }
    build_blocks(c, *k, c.cfg.rows_per_block > 0 ? c.cfg.rows_per_block : 2048);
    if (c.cfg.preconditioner == MHFEM_PREC_ILUT) k->ilu.a0.alloc(c.nnz);
    // level tables: ILU(0), the ISAI factors and SOR's sweeps
    if (c.cfg.preconditioner == MHFEM_PREC_ILU0 || is_isai(c.cfg.preconditioner) || c.cfg.solver == MHFEM_SOLVER_SOR) {
        Ilu0 &I = k->ilu;
        const std::vector<int> dg = build_levels(I, n, c.h_rowptr.data(), c.h_col.data());
        I.lu.alloc(c.nnz);
        if (c.cfg.precision != MHFEM_PRECISION_FP64) I.lu32.alloc(c.nnz);
        if (c.cfg.preconditioner == MHFEM_PREC_ILU0 && ilu_sweep_tables(c, I, dg)) {
            // single-launch triangular solves (k_ilu_sweep)
            I.zmid.alloc(n);
            if (c.cfg.precision != MHFEM_PRECISION_FP64) I.zmid32.alloc(n);
            I.sweep_state.alloc(3);
        }
        if (is_isai(c.cfg.preconditioner)) isai_patterns(c, *k, dg);
        if (c.cfg.preconditioner == MHFEM_PREC_PARILU_ISAI || c.cfg.solver == MHFEM_SOLVER_SOR) {
            I.a0.alloc(c.nnz);
            I.swp.alloc(c.nnz);
        }
    }
    if (c.cfg.preconditioner == MHFEM_PREC_AMG) k->amg = amg_create();
    if (c.cfg.precision == MHFEM_PRECISION_FP64 && c.cfg.face_groups != 0 && k->nseg == 1 && c.unit_diagonal &&
        c.cfg.solver == MHFEM_SOLVER_BICGSTAB &&
        (c.cfg.preconditioner == MHFEM_PREC_DIAGONAL || c.cfg.preconditioner == MHFEM_PREC_NONE)) {
        // the FP64 BiCGStab iteration in the grouped face order: a permuted
        // copy of the matrix, the right-hand side and the iterate translated
        // at the start and the end of each solve
        build_fp32_order(c, *k);
        k->permuted = false;
        k->permuted64 = true;
        k->val64p.alloc(k->diag_base32 + (int64_t)c.F * c.n * (c.n - 1));
        k->bp.alloc(n);
        k->xp.alloc(n);
    }
    if (c.cfg.precision != MHFEM_PRECISION_FP64) {
        // the gather-friendly order needs a plain FP32 iteration: one system,
        // the implicit unit-diagonal Jacobi (or none), the block layout
        const bool reorder = k->nseg == 1 && c.unit_diagonal && c.cfg.face_groups != 0 &&
                             (c.cfg.preconditioner == MHFEM_PREC_DIAGONAL || c.cfg.preconditioner == MHFEM_PREC_NONE);
        if (reorder) build_fp32_order(c, *k);
        k->val32.alloc(k->permuted ? k->diag_base32 + (int64_t)c.F * c.n * (c.n - 1) : c.sell_size);
        for (DBuf<float> *b : {&k->dinv32, &k->b32, &k->x32, &k->r32, &k->r_ast32, &k->p32, &k->Ap32, &k->s32,
                               &k->As32, &k->tmp32, &k->rhs32})
            b->alloc(n);
    }
    return k;
}

// Destroys the captured step graphs once the work queued on their stream
// (replays included) has finished.
static void drop_graphs(Krylov &k)
{
    if (!k.g_first && !k.g_more && !k.g_post) return;
    if (k.g_stream) (void)hipStreamSynchronize(k.g_stream);
    for (hipGraphExec_t *g : {&k.g_first, &k.g_more, &k.g_post})
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

void krylov_destroy(Krylov *k)
{
    if (k) amg_destroy(k->amg);
    if (!k) return;
    drop_graphs(*k);
    if (k->h_sc) (void)hipHostFree(k->h_sc);
    for (int q = 0; q < 2; q++) {
        if (k->h_sc_ahead[q]) (void)hipHostFree(k->h_sc_ahead[q]);
        if (k->ev_ahead[q]) (void)hipEventDestroy(k->ev_ahead[q]);
    }
    delete k;
}

// ---------------------------------------------------------------------------
// host-side driver
namespace {

struct Solve {
    Ctx &c;
    Krylov &k;
    int n;
    unsigned g;
    Seg sg;

    Solve(Ctx &c_, Krylov &k_) : c(c_), k(k_), n(k_.n), g(grid_for(k_.n))
    {
        sg = Seg{k.nseg, k.nblk, k.blk_seg.p, k.blk_row.p, k.seg_blk.p, k.sc.p, k.partials.p, nullptr, k.nblk};
    }

    // The segments a launch covers: M_ACTIVE work goes over sg, which may
    // be narrowed to the blocks of the segments still running (Narrowed);
    // every other mode takes all blocks.
    Seg seg(int mode) const
    {
        if (mode == M_ACTIVE) return sg;
        Seg s = sg;
        s.active = nullptr;
        s.nact = s.nblk;
        return s;
    }
    void widen()
    {
        sg.active = nullptr;
        sg.nact = sg.nblk;
    }

    // Narrowing of the M_ACTIVE launches of one eager solve (bicg_drive)
    // to the blocks of the segments that the latest host copy of the flags
    // (k.h_sc) shows running, once those are at most 9/10 of the work-groups
    // launched now (3/5 and 4/5 were measured 1.6-2 % slower on a 96-instance
    // ensemble; narrowing at every drop about the same): the tail of an ensemble whose segments converge at
    // different iteration counts then launches its few running segments'
    // blocks instead of every block only to see most of them return.
    // A segment halted in that copy does not run again before a restart,
    // which widen()s first; one that halted after the copy is still
    // launched and returns on its own flags.  The list goes to k.active in
    // stream order, after the kernels queued with the previous list; each
    // host list stays alive until the stream has read it (a synchronising
    // flag read, or the end of the solve).  A captured graph keeps the grids of its capture: the graph
    // step replays its graph only while the launches are not narrowed.
    struct Narrowed {
        Solve &S;
        std::vector<std::vector<int>> lists;
        explicit Narrowed(Solve &s) : S(s) {}
        ~Narrowed()
        {
            // an exception can leave copies from the host lists queued
            if (!lists.empty()) (void)hipStreamSynchronize(S.c.stream);
            S.widen();
        }
        // after a synchronising flag read: every copy queued so far is done
        void synchronized() { lists.clear(); }
        void update()
        {
            const Krylov &k = S.k;
            if (k.nseg < 2 || S.any_restart()) return;
            std::vector<int> a;
            for (int s = 0; s < k.nseg; s++) {
                if (S.seg_sc(s)[S_DONE] != 0.0) continue;
                for (int b = k.h_seg_blk[s]; b < k.h_seg_blk[s + 1]; b++) a.push_back(b);
            }
            if (a.empty() || 10 * (long)a.size() > 9 * (long)S.sg.nact) return;
            lists.push_back(std::move(a));
            const std::vector<int> &l = lists.back();
            MHFEM_HIP(hipMemcpyAsync(k.active.p, l.data(), l.size() * sizeof(int), hipMemcpyHostToDevice,
                                     S.c.stream));
            S.sg.active = k.active.p;
            S.sg.nact = (int)l.size();
        }
    };

    // the FP64 BiCGStab solve runs in the grouped face order (k.permuted64)
    bool perm64 = false;

    template <typename T>
    SellView<T> matrix()
    {
        if constexpr (sizeof(T) == 8)
            return perm64 ? SellView<T>{c.F, k.slice_ptr32.p, k.slice_width32.p, k.col32.p, k.val64p.p, k.diag_base32}
                          : SellView<T>{c.F, c.slice_ptr.p, c.slice_width.p, c.sell_col.p, c.sell_val.p, c.diag_base};
        else
            return k.permuted ? SellView<T>{c.F, k.slice_ptr32.p, k.slice_width32.p, k.col32.p, k.val32.p, k.diag_base32}
                              : SellView<T>{c.F, c.slice_ptr.p, c.slice_width.p, c.sell_col.p, k.val32.p, c.diag_base};
    }

    // launches of the matrix kernels, dispatched on the number of equations
    template <typename T, bool DIAG, int ND>
    void sspmv(const T *x, const T *dinv, T *y, const T *z, int mode, const T *w = nullptr)
    {
        SellView<T> A = matrix<T>();
        const bool timed = kernel_timing_begin(c);
        // Threads per work-group, measured at n = 2 on the 64^3 Kuhn mesh:
        // the FP32 product runs best with 512 (one face per thread at 1024
        // rows per work-group: 89-91 us, against 92 us with 256 threads at 512
        // rows and 91-98 us with 1024 threads at 2048 rows); the FP64 product
        // with 256 (186 us at 512 rows; 213 us with 1024 threads at 2048).
        const int fpb = k.rows_per_block / c.n;
        with_n(c.n, [&](auto NC) {
            auto go = [&](auto BSC) {
                constexpr int BS = decltype(BSC)::value;
                hipLaunchKernelGGL((k_sspmv<T, decltype(NC)::value, DIAG, ND, BS>), seg(mode).nact, BS, 0, c.stream,
                                   A, seg(mode), mode, x, dinv, y, z, w);
            };
            if (sizeof(T) == 4 && fpb >= 512) go(std::integral_constant<int, 512>{});
            else go(std::integral_constant<int, KB>{});
        });
        MHFEM_HIP(hipGetLastError());
        if (timed) kernel_timing_end(c);
    }
    template <typename T>
    void sresidual(const T *x, const T *b, T *y, int mode)
    {
        SellView<T> A = matrix<T>();
        with_n(c.n, [&](auto NC) {
            hipLaunchKernelGGL((k_sresidual<T, decltype(NC)::value>), seg(mode).nact, KB, 0, c.stream, A, seg(mode), mode, x, b, y);
        });
        MHFEM_HIP(hipGetLastError());
    }
    template <typename T>
    void spmv_plain(const T *x, T *y)
    {
        SellView<T> A = matrix<T>();
        with_n(c.n, [&](auto NC) {
            hipLaunchKernelGGL((k_spmv_plain<T, decltype(NC)::value>), grid_for(c.F), KB, 0, c.stream, A, x, y);
        });
        MHFEM_HIP(hipGetLastError());
    }
    template <typename T>
    void residual(const T *x, const T *b, T *y)
    {
        SellView<T> A = matrix<T>();
        with_n(c.n, [&](auto NC) {
            hipLaunchKernelGGL((k_residual<T, decltype(NC)::value>), grid_for(c.F), KB, 0, c.stream, A, x, b, y);
        });
        MHFEM_HIP(hipGetLastError());
    }
    void residual_mixed(const double *x, const double *b, float *r32)
    {
        SellView<double> A = matrix<double>();
        with_n(c.n, [&](auto NC) {
            hipLaunchKernelGGL((k_residual_mixed<decltype(NC)::value>), grid_for(c.F), KB, 0, c.stream, A, x, b, r32,
                               k.permuted ? (const int *)k.iperm.p : (const int *)nullptr);
        });
        MHFEM_HIP(hipGetLastError());
    }

    // x64 += d32 (in the FP32 iteration's face order); d32 = 0
    void add_correction(double *x, float *d32)
    {
        if (k.permuted)
            with_n(c.n, [&](auto NC) {
                hipLaunchKernelGGL((k_add_correction_perm<decltype(NC)::value>), grid_for(c.F), KB, 0, c.stream, c.F,
                                   x, d32, (const int *)k.iperm.p);
            });
        else
            hipLaunchKernelGGL(k_add_correction, g, KB, 0, c.stream, n, x, d32);
        MHFEM_HIP(hipGetLastError());
    }

    void set_scalars(double tol, int maxit)
    {
        hipLaunchKernelGGL(k_init_scalars, (k.nseg + 63) / 64, 64, 0, c.stream, k.sc.p, k.nseg, tol, (double)maxit);
        MHFEM_HIP(hipGetLastError());
    }
    void poll()
    {
        MHFEM_HIP(hipMemcpyAsync(k.h_sc, k.sc.p, (size_t)k.nseg * NSLOTS * sizeof(double), hipMemcpyDeviceToHost,
                                 c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    }
    const double *seg_sc(int s) const { return k.h_sc + (long)s * NSLOTS; }
    bool all_done() const
    {
        for (int s = 0; s < k.nseg; s++)
            if (seg_sc(s)[S_DONE] == 0.0) return false;
        return true;
    }
    bool any_restart() const
    {
        for (int s = 0; s < k.nseg; s++)
            if (seg_sc(s)[S_RESTART] != 0.0) return true;
        return false;
    }

    template <int ND>
    void finish(int fin, int mode)
    {
        hipLaunchKernelGGL((k_finalize<ND>), k.nseg, KB, 0, c.stream, sg, fin, mode);
        MHFEM_HIP(hipGetLastError());
    }
    // NF: partial rows the finalize sums (ND, or more when earlier kernels
    // left their partials in the rows past ND)
    template <typename T, int ND, int NF = ND>
    void dots(DotArgs<T, ND> a, int fin, int mode)
    {
        hipLaunchKernelGGL((k_sdots<T, ND>), seg(mode).nact, KB, 0, c.stream, seg(mode), mode, a);
        finish<NF>(fin, mode);
    }
    // FP32 inner solves with the preconditioner fused into the products
    // (Jacobi, none) run the merged BiCGStab iteration.  With a separate
    // preconditioner (ILU(0)) the classic one needed fewer iterations (155
    // against 170-178 per step at 64^3, n = 2).
    template <typename T>
    bool merged() const
    {
        return sizeof(T) == 4 && !separate_prec();
    }

    // The assembled MHFEM rows are scaled so that every diagonal entry is
    // exactly 1 (LinearSystem.h:196-201, BoundaryConditions.hpp:200-203,
    // :375-377): the diagonal preconditioner is then the identity and its
    // scaling vector is not read.
    bool jacobi_scales() const { return c.cfg.preconditioner == MHFEM_PREC_DIAGONAL && !c.unit_diagonal; }

    // ---- preconditioner apply y = M^{-1} x (T = double or float)
    template <typename T>
    void ilu_apply(const T *x, T *y, bool skip_when_halted)
    {
        Ilu0 &I = k.ilu;
        const T *lu;
        if constexpr (sizeof(T) == 8) lu = I.lu.p;
        else lu = I.lu32.p;
        const double *flag = (skip_when_halted && k.nseg == 1) ? k.sc.p : nullptr;
        const int *rp = I.own ? I.frp.p : c.rowptr.p, *col = I.own ? I.fcol.p : c.csr_col.p;
        if (!I.own && I.sweep_state.p) {
            // ILU(0): one launch per triangular solve (k_ilu_sweep)
            T *z;
            if constexpr (sizeof(T) == 8) z = I.zmid.p;
            else z = I.zmid32.p;
            const int nchunks = I.nchunks;
            const IluSweep a{c.F, nchunks, rp, I.diag.p, I.slot.p, I.imp_ptr.p, I.imp_rows.p, I.chunk_ptr.p,
                             I.chunk_face.p, I.chunk_order.p, I.sweep_state.p, I.sweep_state.p + 2, k.sc.p, k.nseg};
            with_n(c.n, [&](auto NC) {
                constexpr int N = decltype(NC)::value;
                hipLaunchKernelGGL((k_ilu_sweep<T, N, false>), nchunks, ILU_WG + ILU_NF, 0, c.stream, a, lu, x, z, y,
                                   flag);
                hipLaunchKernelGGL((k_ilu_sweep<T, N, true>), nchunks, ILU_WG + ILU_NF, 0, c.stream, a, lu,
                                   (const T *)z, y, z, flag);
            });
            MHFEM_HIP(hipGetLastError());
            return;
        }
        for (size_t l = 0; l + 1 < I.h_lptr.size(); l++) {
            const int cnt = I.h_lptr[l + 1] - I.h_lptr[l];
            hipLaunchKernelGGL((k_ilu_lower<T>), (cnt + KB - 1) / KB, KB, 0, c.stream, I.lrows.p + I.h_lptr[l], cnt,
                               rp, col, I.diag.p, lu, x, y, flag);
        }
        for (size_t l = 0; l + 1 < I.h_uptr.size(); l++) {
            const int cnt = I.h_uptr[l + 1] - I.h_uptr[l];
            hipLaunchKernelGGL((k_ilu_upper<T>), (cnt + KB - 1) / KB, KB, 0, c.stream, I.urows.p + I.h_uptr[l], cnt,
                               rp, col, I.diag.p, lu, y, flag);
        }
        MHFEM_HIP(hipGetLastError());
    }
    // one V-cycle (FP64 only; the setup rejects AMG with an FP32 solver)
    template <typename T>
    void amg_apply_t(const T *x, T *y, bool skip_when_halted)
    {
        if constexpr (sizeof(T) == 8) {
            amg_apply(c, *k.amg, x, y, (skip_when_halted && k.nseg == 1) ? k.sc.p : nullptr);
        } else {
            (void)x, (void)y, (void)skip_when_halted;
            throw Error{MHFEM_E_STATE, "the AMG preconditioner works in FP64"};
        }
    }
    // y = M_U (M_L x): two CSR products (FP64 only, as AMG)
    template <typename T>
    void isai_apply_t(const T *x, T *y, bool skip_when_halted)
    {
        if constexpr (sizeof(T) == 8) {
            const Isai &S = k.isai;
            const double *flag = (skip_when_halted && k.nseg == 1) ? k.sc.p : nullptr;
            isai_mv(c, S.lo, x, S.t.p, flag);
            isai_mv(c, S.up, S.t.p, y, flag);
            MHFEM_HIP(hipGetLastError());
        } else {
            (void)x, (void)y, (void)skip_when_halted;
            throw Error{MHFEM_E_STATE, "the ILU-ISAI preconditioner works in FP64"};
        }
    }
    // preconditioners applied by their own kernels after the matrix product
    bool separate_prec() const
    {
        const int p = c.cfg.preconditioner;
        return p == MHFEM_PREC_ILU0 || p == MHFEM_PREC_ILUT || p == MHFEM_PREC_AMG || is_isai(p);
    }
    template <typename T>
    void separate_apply(const T *x, T *y, bool skip_when_halted)
    {
        if (c.cfg.preconditioner == MHFEM_PREC_AMG) amg_apply_t<T>(x, y, skip_when_halted);
        else if (is_isai(c.cfg.preconditioner)) isai_apply_t<T>(x, y, skip_when_halted);
        else ilu_apply<T>(x, y, skip_when_halted);
    }
    template <typename T>
    void prec(const T *x, T *y, const T *dinv, int mode)
    {
        if (separate_prec()) separate_apply<T>(x, y, false);
        else hipLaunchKernelGGL((k_sscale<T>), seg(mode).nact, KB, 0, c.stream, seg(mode), mode, x, jacobi_scales() ? dinv : (const T *)nullptr, y);
        MHFEM_HIP(hipGetLastError());
    }

    // y = M^{-1} A x on the active segments; ND fused dots: 1 -> (y, z),
    // 2 -> (y, z), (y, y), 4 -> (y, z), (y, y), (y, w), (z, w); then `fin`
    template <typename T, int ND, int NF = ND>
    void pmatvec(const T *x, T *y, T *tmp, const T *dinv, const T *z, int fin, const T *w = nullptr)
    {
        if (separate_prec()) {
            sspmv<T, false, 0>(x, (const T *)nullptr, tmp, (const T *)nullptr, M_ACTIVE);
            separate_apply<T>(tmp, y, true);
            if constexpr (ND == 1) dots<T, 1, NF>(DotArgs<T, 1>{{y}, {z}}, fin, M_ACTIVE);
            if constexpr (ND == 2) dots<T, 2>(DotArgs<T, 2>{{y, y}, {z, y}}, fin, M_ACTIVE);
            if constexpr (ND == 4) dots<T, 4>(DotArgs<T, 4>{{y, y, y, z}, {z, y, w, w}}, fin, M_ACTIVE);
        } else {
            if (jacobi_scales())
                sspmv<T, true, ND>(x, dinv, y, z, M_ACTIVE, w);
            else
                sspmv<T, false, ND>(x, dinv, y, z, M_ACTIVE, w);
            if constexpr (ND > 0) finish<NF>(fin, M_ACTIVE);
        }
        MHFEM_HIP(hipGetLastError());
    }
    // r = M^{-1}(b - A x) on the segments selected by mode.  ILU(0) works on
    // every row, so for a partial selection it writes into `scratch` and the
    // selected rows are copied to r.
    template <typename T>
    void presidue(const T *x, const T *b, T *r, T *tmp, const T *dinv, int mode, T *scratch = nullptr)
    {
        sresidual<T>(x, b, tmp, mode);
        if (separate_prec() && mode != M_ALL) {
            if (!scratch) throw Error{MHFEM_E_STATE, "restart with ILU(0), ILU-ISAI or AMG needs a scratch vector"};
            separate_apply<T>(tmp, scratch, false);
            hipLaunchKernelGGL((k_scopy2<T>), seg(mode).nact, KB, 0, c.stream, seg(mode), mode, (const T *)scratch, r, (T *)nullptr);
            MHFEM_HIP(hipGetLastError());
            return;
        }
        prec<T>(tmp, r, dinv, mode);
    }

    struct Vecs {
        void *r, *r_ast, *p, *Ap, *s, *As, *tmp, *tmp2, *dinv;
    };

    // ---- BiCGStab pieces (left preconditioned, restart on rho breakdown,
    // stop when ||r_k|| / ||M^{-1} b|| < tol), shared by the eager solve and
    // the captured step graphs
    template <typename T>
    void bicg_start(const T *b, T *x, const Vecs &v, double tol, int maxit)
    {
        set_scalars(tol, maxit);
        prec<T>(b, (T *)v.tmp2, (const T *)v.dinv, M_ALL);
        dots<T, 1>(DotArgs<T, 1>{{(T *)v.tmp2}, {(T *)v.tmp2}}, FIN_BNORM, M_ALL);
        bicg_init<T>(b, x, v, M_ALL);
    }
    template <typename T>
    void bicg_init(const T *b, T *x, const Vecs &v, int mode)
    {
        T *r = (T *)v.r;
        presidue<T>(x, b, r, (T *)v.tmp, (const T *)v.dinv, mode, (T *)v.tmp2);
        hipLaunchKernelGGL((k_scopy2<T>), seg(mode).nact, KB, 0, c.stream, seg(mode), mode, r, (T *)v.p, (T *)v.r_ast);
        MHFEM_HIP(hipGetLastError());
        if (merged<T>())  // (r, r) also in partial row 1, for the first merged alpha
            dots<T, 2>(DotArgs<T, 2>{{r, r}, {(T *)v.r_ast, r}}, FIN_INIT, mode);
        else
            dots<T, 1>(DotArgs<T, 1>{{r}, {(T *)v.r_ast}}, FIN_INIT, mode);
    }
    template <typename T>
    void bicg_iterations(T *x, const Vecs &v, int count)
    {
        T *r = (T *)v.r, *r_ast = (T *)v.r_ast, *p = (T *)v.p, *Ap = (T *)v.Ap, *s = (T *)v.s, *As = (T *)v.As,
          *tmp = (T *)v.tmp;
        const T *dinv = (const T *)v.dinv;
        for (int q = 0; q < count; q++) {
            if (merged<T>()) {
                // alpha with the previous iteration's convergence test;
                // omega and beta from the second product's four dots; x, r
                // and p in one pass
                pmatvec<T, 1, 2>(p, Ap, tmp, dinv, r_ast, FIN_ALPHA_MERGED);
                hipLaunchKernelGGL((k_bicg_s<T>), sg.nact, KB, 0, c.stream, sg, r, Ap, s);
                pmatvec<T, 4>(s, As, tmp, dinv, s, FIN_OMEGA_MERGED, r_ast);
                hipLaunchKernelGGL((k_bicg_xrp<T>), sg.nact, KB, 0, c.stream, sg, x, p, s, As, r, Ap);
            } else {
                pmatvec<T, 1>(p, Ap, tmp, dinv, r_ast, FIN_ALPHA);
                hipLaunchKernelGGL((k_bicg_s<T>), sg.nact, KB, 0, c.stream, sg, r, Ap, s);
                pmatvec<T, 2>(s, As, tmp, dinv, s, FIN_OMEGA);
                hipLaunchKernelGGL((k_bicg_xr<T>), sg.nact, KB, 0, c.stream, sg, x, p, s, As, r, r_ast);
                finish<2>(FIN_UPDATE, M_ACTIVE);
                hipLaunchKernelGGL((k_bicg_p<T>), sg.nact, KB, 0, c.stream, sg, r, p, Ap);
            }
            MHFEM_HIP(hipGetLastError());
        }
    }
    // eager solve; returns the largest iteration count over the segments
    template <typename T>
    int bicgstab(const T *b, T *x, const Vecs &v, double tol, int maxit)
    {
        bicg_start<T>(b, x, v, tol, maxit);
        return bicg_drive<T>(b, x, v, [&]() { bicg_iterations<T>(x, v, std::max(1, c.cfg.check_every)); });
    }
    // poll the flags, restart segments that need it, queue more iterations
    template <typename T, typename More>
    // Poll, restart the segments that need it, queue more iterations.  While
    // the residual drop of the last batch says that more than two further
    // batches are needed, one batch is kept queued ahead of the flags the host
    // is reading (an asynchronous copy into pinned memory and an event), so
    // the GPU does not wait for the host between batches.  Near convergence
    // the host waits for each batch instead: a batch queued after convergence
    // would cost its launches (the kernels of a finished segment return at
    // once) and short solves would lose more than they gain.
    int bicg_drive(const T *b, T *x, const Vecs &v, More &&more)
    {
        Narrowed narrowed(*this);
        const size_t bytes = (size_t)k.nseg * NSLOTS * sizeof(double);
        auto snap = [&](int q) {
            MHFEM_HIP(hipMemcpyAsync(k.h_sc_ahead[q], k.sc.p, bytes, hipMemcpyDeviceToHost, c.stream));
            MHFEM_HIP(hipEventRecord(k.ev_ahead[q], c.stream));
        };
        auto take = [&](int q) {
            MHFEM_HIP(hipEventSynchronize(k.ev_ahead[q]));
            std::memcpy(k.h_sc, k.h_sc_ahead[q], bytes);
        };
        // residuals one batch earlier; 0 = unknown (no batch measured yet)
        std::vector<double> prev(k.nseg, 0.0);
        auto remember = [&]() {
            for (int s = 0; s < k.nseg; s++) prev[s] = seg_sc(s)[S_RESID];
        };
        // batches the slowest running segment still needs, by geometric
        // extrapolation of its last batch (0 while it is unknown)
        auto batches_left = [&]() {
            double worst = 0.0;
            for (int s = 0; s < k.nseg; s++) {
                const double *f = seg_sc(s);
                if (f[S_DONE] != 0.0 || f[S_RESTART] != 0.0) continue;
                const double r = f[S_RESID], t = f[S_TOL], p = prev[s];
                if (!(r > t) || !(p > 0.0)) continue;
                worst = std::max(worst, p > r ? std::log(t / r) / std::log(r / p) : 1e9);
            }
            return worst;
        };
        poll();
        for (;;) {
            if (all_done()) break;
            if (any_restart()) {
                widen();
                bicg_init<T>(b, x, v, M_RESTART);
                std::fill(prev.begin(), prev.end(), 0.0);
                poll();
                continue;
            }
            more();
            if (batches_left() > 2.0) {
                snap(0);
                for (int q = 0;; q ^= 1) {
                    more();
                    snap(q ^ 1);
                    remember();
                    take(q);
                    narrowed.update();
                    if (all_done() || any_restart() || !(batches_left() > 2.0)) break;
                }
            }
            remember();
            poll();  // the batch still queued ends; the latest flags decide
            narrowed.synchronized();
            narrowed.update();
        }
        int it = 0;
        for (int s = 0; s < k.nseg; s++) it = std::max(it, (int)seg_sc(s)[S_IT]);
        return it;
    }

    // GMRES(m), FP64, one segment
    int gmres(const double *b, double *x, double tol, int maxit, int m, double *resid_out);
    int cg(const double *b, double *x, double tol, int maxit, double *resid_out);
    int tfqmr(const double *b, double *x, double tol, int maxit, double *resid_out);
    int richardson(const double *b, double *x, double tol, int maxit, double *resid_out);
    int bicgstabl(const double *b, double *x, double tol, int maxit, int ell, double *resid_out);
    int idrs(const double *b, double *x, double tol, int maxit, int s, double *resid_out);
    int sor(const double *b, double *x, double tol, int maxit, double omega, double *resid_out);
    void pmv(const double *in, double *out);  // out = M^{-1} A in (through k.tmp)
    double *ext(size_t vectors);             // k.ext grown to hold `vectors` vectors
    // host-driven pieces of those three: y = a u + b v (y may be u or v),
    // dot products read back to the host, device copies
    void lin2(double a, const double *u, double b, const double *v, double *y);
    double dot1(const double *x, const double *y);
    void dot2(const double *x0, const double *y0, const double *x1, const double *y1, double *d0, double *d1);
    void copy(const double *x, double *y);
};

// ---- GMRES pieces: partial dots of w with V_0..V_{j1-1}, one row of partials per vector
__global__ void __launch_bounds__(KB) k_gs_dots(int n, int j1, const double *__restrict__ V, const double *__restrict__ w,
                                                double *__restrict__ partials)
{
    __shared__ double sm[KB / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int kk = 0; kk < j1; kk++) {
        double a = 0;
        for (int i = blockIdx.x * KB + tid; i < n; i += gridDim.x * KB) a += V[(long)kk * n + i] * w[i];
        a = wave_sum(a);
        if (lane == 0) sm[wid] = a;
        __syncthreads();
        if (tid == 0) {
            double s = sm[0];
            for (int q = 1; q < KB / 64; q++) s += sm[q];
            partials[(long)kk * gridDim.x + blockIdx.x] = s;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(KB) k_reduce_rows(int rows, int cols, const double *__restrict__ partials,
                                                    double *__restrict__ out)
{
    __shared__ double sm[KB / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int rr = blockIdx.x;
    if (rr >= rows) return;
    double a = 0;
    for (int q = tid; q < cols; q += KB) a += partials[(long)rr * cols + q];
    a = wave_sum(a);
    if (lane == 0) sm[wid] = a;
    __syncthreads();
    if (tid == 0) {
        double s = sm[0];
        for (int q = 1; q < KB / 64; q++) s += sm[q];
        out[rr] = s;
    }
}

__global__ void __launch_bounds__(KB) k_gs_update(int n, int j1, const double *__restrict__ V, const double *__restrict__ h,
                                                  double *__restrict__ w)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        double s = w[i];
        for (int kk = 0; kk < j1; kk++) s -= h[kk] * V[(long)kk * n + i];
        w[i] = s;
    }
}

__global__ void __launch_bounds__(KB) k_scale_into(int n, const double *__restrict__ w, double inv, double *__restrict__ v)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) v[i] = w[i] * inv;
}

__global__ void __launch_bounds__(KB) k_combine(int n, int j, const double *__restrict__ V, const double *__restrict__ y,
                                                double *__restrict__ x)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        double s = 0;
        for (int kk = 0; kk < j; kk++) s += y[kk] * V[(long)kk * n + i];
        x[i] += s;
    }
}

int Solve::gmres(const double *b, double *x, double tol, int maxit, int m, double *resid_out)
{
    if (m < 1) m = 30;
    if (k.V_m != m) {
        k.V.alloc((size_t)(m + 1) * n);
        k.V_m = m;
    }
    double *V = k.V.p, *w = k.tmp2.p, *tmp = k.tmp.p;
    const double *dinv = k.dinv.p;
    DBuf<double> hbuf;
    hbuf.alloc(2 * (m + 2));
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), h(m + 2), h2(m + 2);
    double *hy = hbuf.p + (m + 2);
    auto norm2 = [&](const double *v) {
        dots<double, 1>(DotArgs<double, 1>{{v}, {v}}, FIN_NONE, M_ALL);
        poll();
        return std::sqrt(k.h_sc[S_D0]);
    };
    set_scalars(tol, maxit);
    prec<double>(b, w, dinv, M_ALL);
    double b_norm = norm2(w);
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    presidue<double>(x, b, V, tmp, dinv, M_ALL);
    double beta = norm2(V);
    double resid = beta / b_norm;
    const unsigned gg = std::min<unsigned>(g, 1024);
    DBuf<double> gpart;  // partial dots of the Gram-Schmidt step: (m+1) rows of gg
    gpart.alloc((size_t)(m + 1) * gg);
    while (resid >= tol && it < maxit && !std::isnan(resid)) {
        hipLaunchKernelGGL(k_scale_into, g, KB, 0, c.stream, n, V, 1.0 / beta, V);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int j;
        for (j = 0; j < m && it < maxit; j++) {
            pmatvec<double, 0>(V + (long)j * n, w, tmp, dinv, nullptr, 0);
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL(k_gs_dots, gg, KB, 0, c.stream, n, j + 1, V, w, gpart.p);
                hipLaunchKernelGGL(k_reduce_rows, j + 1, KB, 0, c.stream, j + 1, (int)gg, gpart.p, hbuf.p);
                hipLaunchKernelGGL(k_gs_update, g, KB, 0, c.stream, n, j + 1, V, hbuf.p, w);
                MHFEM_HIP(hipGetLastError());
                MHFEM_HIP(hipMemcpyAsync(h2.data(), hbuf.p, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
                MHFEM_HIP(hipStreamSynchronize(c.stream));
                for (int q = 0; q <= j; q++) h[q] = pass ? h[q] + h2[q] : h2[q];
            }
            h[j + 1] = norm2(w);
            if (h[j + 1] != 0.0)
                hipLaunchKernelGGL(k_scale_into, g, KB, 0, c.stream, n, w, 1.0 / h[j + 1], V + (long)(j + 1) * n);
            for (int q = 0; q < j; q++) {
                const double t = cs[q] * h[q] + sn[q] * h[q + 1];
                h[q + 1] = -sn[q] * h[q] + cs[q] * h[q + 1];
                h[q] = t;
            }
            const double den = std::sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
            cs[j] = den > 0 ? h[j] / den : 1.0;
            sn[j] = den > 0 ? h[j + 1] / den : 0.0;
            h[j] = cs[j] * h[j] + sn[j] * h[j + 1];
            h[j + 1] = 0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            for (int q = 0; q <= j; q++) H[(size_t)q * m + j] = h[q];
            it++;
            resid = std::fabs(gv[j + 1]) / b_norm;
            if (resid < tol) {
                j++;
                break;
            }
        }
        for (int q = j - 1; q >= 0; q--) {
            double s = gv[q];
            for (int t = q + 1; t < j; t++) s -= H[(size_t)q * m + t] * y[t];
            y[q] = s / H[(size_t)q * m + q];
        }
        MHFEM_HIP(hipMemcpyAsync(hy, y.data(), j * sizeof(double), hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_combine, g, KB, 0, c.stream, n, j, V, hy, x);
        MHFEM_HIP(hipGetLastError());
        presidue<double>(x, b, V, tmp, dinv, M_ALL);
        beta = norm2(V);
        resid = beta / b_norm;
    }
    MHFEM_HIP(hipStreamSynchronize(c.stream));
    *resid_out = resid;
    return it;
}

// ---------------------------------------------------------------------------
// CG, TFQMR and the Jacobi/Richardson iteration, FP64, the other solvers of
// the reference's "linear-solver" option (configSetup.h:31-37).  Their scalar
// recurrences run on the host, one read-back per dot product; the vectors stay
// in HBM.  Each stops on ||M^{-1}(b - A x)|| / ||M^{-1} b|| < tol like the
// BiCGStab and GMRES here.

__global__ void __launch_bounds__(KB) k_lin2(int n, double a, const double *u, double b, const double *v, double *y)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) y[i] = a * u[i] + b * v[i];
}

void Solve::lin2(double a, const double *u, double b, const double *v, double *y)
{
    hipLaunchKernelGGL(k_lin2, g, KB, 0, c.stream, n, a, u, b, v, y);
    MHFEM_HIP(hipGetLastError());
}
double Solve::dot1(const double *x, const double *y)
{
    dots<double, 1>(DotArgs<double, 1>{{x}, {y}}, FIN_NONE, M_ALL);
    poll();
    return k.h_sc[S_D0];
}
void Solve::dot2(const double *x0, const double *y0, const double *x1, const double *y1, double *d0, double *d1)
{
    dots<double, 2>(DotArgs<double, 2>{{x0, x1}, {y0, y1}}, FIN_NONE, M_ALL);
    poll();
    *d0 = k.h_sc[S_D0];
    *d1 = k.h_sc[S_D1];
}
void Solve::copy(const double *x, double *y)
{
    MHFEM_HIP(hipMemcpyAsync(y, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
}

// Preconditioned CG (Hestenes & Stiefel 1952), z = M^{-1} r; meant for
// symmetric positive definite systems (TNL::Solvers::Linear::CG)
int Solve::cg(const double *b, double *x, double tol, int maxit, double *resid_out)
{
    double *r = k.r.p, *z = k.s.p, *p = k.p.p, *q = k.Ap.p;
    const double *dinv = k.dinv.p;
    set_scalars(tol, maxit);
    prec<double>(b, z, dinv, M_ALL);
    double b_norm = std::sqrt(dot1(z, z));
    if (b_norm == 0.0) b_norm = 1.0;
    sresidual<double>(x, b, r, M_ALL);
    prec<double>(r, z, dinv, M_ALL);
    copy(z, p);
    double rho, zz;
    dot2(r, z, z, z, &rho, &zz);
    double resid = std::sqrt(zz) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && std::isfinite(resid)) {
        spmv_plain<double>(p, q);
        const double alpha = rho / dot1(p, q);
        lin2(1.0, x, alpha, p, x);
        lin2(1.0, r, -alpha, q, r);
        prec<double>(r, z, dinv, M_ALL);
        double rho_new;
        dot2(r, z, z, z, &rho_new, &zz);
        resid = std::sqrt(zz) / b_norm;
        const double beta = rho_new / rho;
        rho = rho_new;
        lin2(1.0, z, beta, p, p);
        it++;
    }
    MHFEM_HIP(hipStreamSynchronize(c.stream));
    *resid_out = resid;
    return it;
}

// TFQMR (Freund 1993) on M^{-1} A x = M^{-1} b in the two-half-step form
// (Kelley 1995), shadow vector r_0 (TNL::Solvers::Linear::TFQMR).  One
// iteration is a pair of half steps.  When the quasi-residual bound
// tau sqrt(m + 1) falls below tol ||M^{-1} b|| the true preconditioned
// residual is computed and must be below tol too.  A breakdown, (r_0, v) = 0
// or rho = 0, which the identity rows of Dirichlet faces do reach, restarts
// the recurrence from the current x.
double *Solve::ext(size_t vectors)
{
    if (k.ext.n < vectors * n) k.ext.alloc(vectors * n);
    return k.ext.p;
}
void Solve::pmv(const double *in, double *out)
{
    spmv_plain<double>(in, k.tmp.p);
    prec<double>(k.tmp.p, out, k.dinv.p, M_ALL);
}

int Solve::tfqmr(const double *b, double *x, double tol, int maxit, double *resid_out)
{
    double *e = ext(2);
    double *rt = k.r_ast.p, *w = k.r.p, *y1 = k.p.p, *y2 = k.s.p, *u1 = k.Ap.p, *u2 = k.As.p, *v = k.tmp2.p,
           *d = e, *r = e + n, *tmp = k.tmp.p;
    const double *dinv = k.dinv.p;
    set_scalars(tol, maxit);
    prec<double>(b, r, dinv, M_ALL);
    double b_norm = std::sqrt(dot1(r, r));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    bool converged = false, restart = true;
    double tau = 0, theta = 0, eta = 0, rho = 0, resid = 0;
    while (restart) {
        restart = false;
        presidue<double>(x, b, rt, tmp, dinv, M_ALL);
        copy(rt, w);
        copy(rt, y1);
        MHFEM_HIP(hipMemsetAsync(d, 0, (size_t)n * sizeof(double), c.stream));
        pmv(y1, v);
        copy(v, u1);
        tau = std::sqrt(dot1(rt, rt));
        theta = eta = 0;
        rho = tau * tau;
        resid = tau / b_norm;
        if (resid < tol) converged = true;
        while (!converged && it < maxit && std::isfinite(resid)) {
            it++;
            const double sigma = dot1(rt, v), alpha = rho / sigma;
            if (!(std::fabs(sigma) > 0.0) || !std::isfinite(alpha)) {
                restart = true;
                break;
            }
            lin2(1.0, y1, -alpha, v, y2);
            pmv(y2, u2);
            for (int j = 1; j <= 2 && !converged; j++) {
                const double *u = j == 1 ? u1 : u2, *y = j == 1 ? y1 : y2;
                const int m = 2 * it - 2 + j;
                const double f = theta * theta * eta / alpha;
                lin2(1.0, w, -alpha, u, w);
                lin2(1.0, y, f, d, d);
                theta = std::sqrt(dot1(w, w)) / tau;
                const double cth = 1.0 / std::sqrt(1.0 + theta * theta);
                tau = tau * theta * cth;
                eta = cth * cth * alpha;
                lin2(1.0, x, eta, d, x);
                if (tau * std::sqrt((double)m + 1.0) < tol * b_norm) {
                    presidue<double>(x, b, r, tmp, dinv, M_ALL);
                    resid = std::sqrt(dot1(r, r)) / b_norm;
                    converged = resid < tol;
                }
            }
            if (converged) break;
            const double rho_new = dot1(rt, w), beta = rho_new / rho;
            if (!(std::fabs(rho_new) > 0.0) || !std::isfinite(beta)) {
                restart = true;
                break;
            }
            rho = rho_new;
            lin2(1.0, w, beta, y2, y1);
            pmv(y1, u1);
            lin2(1.0, u2, beta, v, v);  // v = u1 + beta (u2 + beta v)
            lin2(1.0, u1, beta, v, v);
        }
    }
    presidue<double>(x, b, r, tmp, dinv, M_ALL);
    resid = std::sqrt(dot1(r, r)) / b_norm;
    *resid_out = resid;
    return it;
}

// BiCGStab(l) (Sleijpen & Fokkema 1993) on M^{-1} A x = M^{-1} b: l BiCG steps,
// then a minimal-residual polynomial of degree l over the l + 1 residuals
// (modified Gram-Schmidt) (TNL::Solvers::Linear::BICGStabL).  One iteration is
// one BiCG step (two operator applications); convergence is checked after each
// cycle.  A breakdown (rho0 = 0, (u_{j+1}, r~) = 0, sigma_j = 0) restarts from
// the current x.  l = 1 is BiCGStab.
int Solve::bicgstabl(const double *b, double *x, double tol, int maxit, int ell, double *resid_out)
{
    ell = std::max(1, std::min(ell, 8));
    double *R = ext((size_t)2 * (ell + 1));
    double *U = R + (size_t)(ell + 1) * n, *rt = k.r_ast.p;
    auto Rv = [&](int j) { return R + (size_t)j * n; };
    auto Uv = [&](int j) { return U + (size_t)j * n; };
    double tau[9][9] = {}, sigma[9] = {}, gp[9] = {}, gg[9] = {}, gpp[9] = {};
    set_scalars(tol, maxit);
    prec<double>(b, rt, k.dinv.p, M_ALL);
    double b_norm = std::sqrt(dot1(rt, rt));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    double resid = 0, rho0 = 1, alpha = 0, omega = 1;
    bool restart = true;
    while (restart) {
        restart = false;
        presidue<double>(x, b, R, k.tmp.p, k.dinv.p, M_ALL);
        copy(R, rt);
        MHFEM_HIP(hipMemsetAsync(U, 0, (size_t)n * sizeof(double), c.stream));
        rho0 = 1, alpha = 0, omega = 1;
        resid = std::sqrt(dot1(R, R)) / b_norm;
        while (!restart && resid >= tol && it < maxit && std::isfinite(resid)) {
            rho0 = -omega * rho0;
            for (int j = 0; j < ell && !restart; j++) {
                it++;
                const double rho1 = dot1(Rv(j), rt), beta = alpha * rho1 / rho0;
                if (!(std::fabs(rho0) > 0.0) || !std::isfinite(beta)) {
                    restart = true;
                    break;
                }
                rho0 = rho1;
                for (int i = 0; i <= j; i++) lin2(1.0, Rv(i), -beta, Uv(i), Uv(i));
                pmv(Uv(j), Uv(j + 1));
                const double gam = dot1(Uv(j + 1), rt);
                alpha = rho0 / gam;
                if (!(std::fabs(gam) > 0.0) || !std::isfinite(alpha)) {
                    restart = true;
                    break;
                }
                for (int i = 0; i <= j; i++) lin2(1.0, Rv(i), -alpha, Uv(i + 1), Rv(i));
                pmv(Rv(j), Rv(j + 1));
                lin2(1.0, x, alpha, Uv(0), x);
            }
            if (restart) break;
            for (int j = 1; j <= ell && !restart; j++) {
                for (int i = 1; i < j; i++) {
                    tau[i][j] = dot1(Rv(j), Rv(i)) / sigma[i];
                    lin2(1.0, Rv(j), -tau[i][j], Rv(i), Rv(j));
                }
                sigma[j] = dot1(Rv(j), Rv(j));
                if (!(sigma[j] > 0.0)) restart = true;
                else gp[j] = dot1(Rv(0), Rv(j)) / sigma[j];
            }
            if (restart) break;
            gg[ell] = gp[ell];
            omega = gg[ell];
            for (int j = ell - 1; j >= 1; j--) {
                double t = gp[j];
                for (int i = j + 1; i <= ell; i++) t -= tau[j][i] * gg[i];
                gg[j] = t;
            }
            for (int j = 1; j < ell; j++) {
                double t = gg[j + 1];
                for (int i = j + 1; i < ell; i++) t += tau[j][i] * gg[i + 1];
                gpp[j] = t;
            }
            lin2(1.0, x, gg[1], Rv(0), x);
            lin2(1.0, Rv(0), -gp[ell], Rv(ell), Rv(0));
            lin2(1.0, Uv(0), -gg[ell], Uv(ell), Uv(0));
            for (int j = 1; j < ell; j++) {
                lin2(1.0, Uv(0), -gg[j], Uv(j), Uv(0));
                lin2(1.0, x, gpp[j], Rv(j), x);
                lin2(1.0, Rv(0), -gp[j], Rv(j), Rv(0));
            }
            resid = std::sqrt(dot1(Rv(0), Rv(0))) / b_norm;
        }
    }
    *resid_out = resid;
    return it;
}

// entry (i, k) of IDR(s)'s shadow space before orthonormalisation: an integer
// hash mapped to [-1, 1) with 24 bits, exact in any precision
__global__ void __launch_bounds__(KB) k_idr_shadow(int n, int k, double *__restrict__ p)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        uint32_t h = (uint32_t)i * 0x9E3779B1u + (uint32_t)(k + 1) * 0x85EBCA77u;
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
        h *= 0x297A2D39u;
        h ^= h >> 15;
        p[i] = (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
    }
}

// IDR(s) with biorthogonalisation (van Gijzen & Sonneveld 2011, Alg. 913) on
// M^{-1} A x = M^{-1} b, omega by minimal residual with the angle correction
// 0.7 (TNL::Solvers::Linear::IDRs).  One iteration is one operator
// application (s + 1 per cycle), convergence checked after each.  The shadow
// space is the orthonormalised hash matrix above.  A breakdown (M_kk = 0,
// t = 0) restarts from the current x.
int Solve::idrs(const double *b, double *x, double tol, int maxit, int s, double *resid_out)
{
    s = std::max(1, std::min(s, 8));
    double *P = ext((size_t)3 * s + 2);
    double *G = P + (size_t)s * n, *U = G + (size_t)s * n, *v = U + (size_t)s * n, *t = v + n, *r = k.r.p;
    auto Pv = [&](int i) { return P + (size_t)i * n; };
    auto Gv = [&](int i) { return G + (size_t)i * n; };
    auto Uv = [&](int i) { return U + (size_t)i * n; };
    double Mm[8][8], f[8], cc[8];
    for (int q = 0; q < s; q++) {
        hipLaunchKernelGGL(k_idr_shadow, g, KB, 0, c.stream, n, q, Pv(q));
        MHFEM_HIP(hipGetLastError());
        for (int i = 0; i < q; i++) lin2(1.0, Pv(q), -dot1(Pv(q), Pv(i)), Pv(i), Pv(q));
        lin2(1.0 / std::sqrt(dot1(Pv(q), Pv(q))), Pv(q), 0.0, Pv(q), Pv(q));
    }
    set_scalars(tol, maxit);
    prec<double>(b, r, k.dinv.p, M_ALL);
    double b_norm = std::sqrt(dot1(r, r));
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    double resid = 0, om = 1;
    bool restart = true;
    while (restart) {
        restart = false;
        presidue<double>(x, b, r, k.tmp.p, k.dinv.p, M_ALL);
        MHFEM_HIP(hipMemsetAsync(G, 0, (size_t)2 * s * n * sizeof(double), c.stream));
        for (int i = 0; i < s; i++)
            for (int j = 0; j < s; j++) Mm[i][j] = i == j;
        om = 1.0;
        resid = std::sqrt(dot1(r, r)) / b_norm;
        while (!restart && resid >= tol && it < maxit && std::isfinite(resid)) {
            for (int i = 0; i < s; i++) f[i] = dot1(Pv(i), r);
            for (int kk = 0; kk < s && resid >= tol && it < maxit; kk++) {
                for (int i = kk; i < s; i++) {  // c = M(k:s, k:s) \ f(k:s), M lower triangular
                    double acc = f[i];
                    for (int j = kk; j < i; j++) acc -= Mm[i][j] * cc[j];
                    cc[i] = acc / Mm[i][i];
                }
                copy(r, v);
                for (int i = kk; i < s; i++) lin2(1.0, v, -cc[i], Gv(i), v);
                lin2(om, v, 0.0, v, v);  // u_k = om v + sum_i c_i u_i (old u_k included)
                for (int i = kk; i < s; i++) lin2(1.0, v, cc[i], Uv(i), v);
                copy(v, Uv(kk));
                pmv(Uv(kk), Gv(kk));
                it++;
                for (int i = 0; i < kk; i++) {
                    const double al = dot1(Pv(i), Gv(kk)) / Mm[i][i];
                    lin2(1.0, Gv(kk), -al, Gv(i), Gv(kk));
                    lin2(1.0, Uv(kk), -al, Uv(i), Uv(kk));
                }
                for (int i = kk; i < s; i++) Mm[i][kk] = dot1(Pv(i), Gv(kk));
                if (!(std::fabs(Mm[kk][kk]) > 0.0)) {
                    restart = true;
                    break;
                }
                const double beta = f[kk] / Mm[kk][kk];
                lin2(1.0, r, -beta, Gv(kk), r);
                lin2(1.0, x, beta, Uv(kk), x);
                resid = std::sqrt(dot1(r, r)) / b_norm;
                for (int i = kk + 1; i < s; i++) f[i] -= beta * Mm[i][kk];
            }
            if (restart || resid < tol || it >= maxit || !std::isfinite(resid)) break;
            copy(r, v);
            pmv(v, t);
            it++;
            double tt, tr;
            dot2(t, t, t, r, &tt, &tr);
            if (!(tt > 0.0)) {
                restart = true;
                break;
            }
            om = tr / tt;
            const double rho = std::fabs(tr) / (std::sqrt(tt) * std::sqrt(dot1(r, r)));
            if (rho < 0.7) om *= 0.7 / rho;
            lin2(1.0, r, -om, t, r);
            lin2(1.0, x, om, v, x);
            resid = std::sqrt(dot1(r, r)) / b_norm;
        }
    }
    *resid_out = resid;
    return it;
}

// one level of an SOR sweep: the rows of a level depend below the diagonal
// only on rows of earlier levels, which already hold this sweep's values in
// xn; right of the diagonal the previous sweep's values (xo) are used, as in
// the sequential sweep
__global__ void __launch_bounds__(KB) k_sor_level(const int *__restrict__ rows, int count, const int *__restrict__ rp,
                                                  const int *__restrict__ col, const int *__restrict__ dg,
                                                  const double *__restrict__ a, const double *__restrict__ b,
                                                  const double *__restrict__ xo, double *__restrict__ xn, double omega)
{
    const int t = blockIdx.x * KB + threadIdx.x;
    if (t >= count) return;
    const int r = rows[t];
    double acc = b[r];
    for (int q = rp[r]; q < dg[r]; q++) acc -= a[q] * xn[col[q]];
    for (int q = dg[r] + 1; q < rp[r + 1]; q++) acc -= a[q] * xo[col[q]];
    xn[r] = (1.0 - omega) * xo[r] + omega * acc / a[dg[r]];
}

// SOR (TNL::Solvers::Linear::SOR): forward sweeps in row order, relaxation
// omega, the sequential sweep's result computed level by level (the ILU(0)
// level tables of the lower triangle).  The preconditioner is not used; the
// stopping test is ||b - A x|| / ||b|| < tol after each sweep (one iteration).
int Solve::sor(const double *b, double *x, double tol, int maxit, double omega, double *resid_out)
{
    Ilu0 &I = k.ilu;
    double *xo = k.tmp2.p, *r = k.r.p;
    hipLaunchKernelGGL((k_gather<double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, c.sell_val.p, c.csr_to_sell.p,
                       I.a0.p);
    MHFEM_HIP(hipGetLastError());
    set_scalars(tol, maxit);
    double b_norm = std::sqrt(dot1(b, b));
    if (b_norm == 0.0) b_norm = 1.0;
    sresidual<double>(x, b, r, M_ALL);
    double resid = std::sqrt(dot1(r, r)) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && std::isfinite(resid)) {
        copy(x, xo);
        for (size_t l = 0; l + 1 < I.h_lptr.size(); l++) {
            const int cnt = I.h_lptr[l + 1] - I.h_lptr[l];
            hipLaunchKernelGGL(k_sor_level, (cnt + KB - 1) / KB, KB, 0, c.stream, I.lrows.p + I.h_lptr[l], cnt,
                               c.rowptr.p, c.csr_col.p, I.diag.p, I.a0.p, b, xo, x, omega);
        }
        MHFEM_HIP(hipGetLastError());
        sresidual<double>(x, b, r, M_ALL);
        resid = std::sqrt(dot1(r, r)) / b_norm;
        it++;
    }
    *resid_out = resid;
    return it;
}

// Stationary iteration x += M^{-1}(b - A x) (relaxation 1): with the diagonal
// preconditioner, or the unit diagonal the MHFEM rows are scaled to, it is the
// Jacobi method (TNL::Solvers::Linear::Jacobi, omega = 1)
int Solve::richardson(const double *b, double *x, double tol, int maxit, double *resid_out)
{
    double *r = k.r.p, *tmp = k.tmp.p;
    const double *dinv = k.dinv.p;
    set_scalars(tol, maxit);
    prec<double>(b, r, dinv, M_ALL);
    double b_norm = std::sqrt(dot1(r, r));
    if (b_norm == 0.0) b_norm = 1.0;
    presidue<double>(x, b, r, tmp, dinv, M_ALL);
    double resid = std::sqrt(dot1(r, r)) / b_norm;
    int it = 0;
    while (resid >= tol && it < maxit && std::isfinite(resid)) {
        lin2(1.0, x, 1.0, r, x);
        presidue<double>(x, b, r, tmp, dinv, M_ALL);
        resid = std::sqrt(dot1(r, r)) / b_norm;
        it++;
    }
    *resid_out = resid;
    return it;
}

}  // namespace

// ---------------------------------------------------------------------------
// Preconditioner update (Solver.hpp:1237-1242): every step, from the freshly
// assembled matrix.
// ILU(0) of the current matrix into k.ilu.lu (CSR order), level by level
static void ilu_factor(Ctx &c, Krylov &k)
{
    Ilu0 &I = k.ilu;
    hipLaunchKernelGGL((k_gather<double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, c.sell_val.p, c.csr_to_sell.p,
                       I.lu.p);
    if (I.factor_slots == 0) {
        int longest = 0;
        for (int r = 0; r < k.n; r++) longest = std::max(longest, c.h_rowptr[r + 1] - c.h_rowptr[r]);
        I.factor_slots = longest <= 4 * ILU_FG ? (longest + ILU_FG - 1) / ILU_FG : -1;
    }
    for (size_t l = 0; l + 1 < I.h_lptr.size(); l++) {
        const int cnt = I.h_lptr[l + 1] - I.h_lptr[l];
        const int *rows = I.lrows.p + I.h_lptr[l];
        const unsigned groups = (unsigned)(((long)cnt * ILU_FG + KB - 1) / KB);
        switch (I.factor_slots) {
        case 1:
            hipLaunchKernelGGL((k_ilu_factor_rows<1>), groups, KB, 0, c.stream, rows, cnt, c.rowptr.p, c.csr_col.p,
                               I.diag.p, I.lu.p);
            break;
        case 2:
            hipLaunchKernelGGL((k_ilu_factor_rows<2>), groups, KB, 0, c.stream, rows, cnt, c.rowptr.p, c.csr_col.p,
                               I.diag.p, I.lu.p);
            break;
        case 3:
            hipLaunchKernelGGL((k_ilu_factor_rows<3>), groups, KB, 0, c.stream, rows, cnt, c.rowptr.p, c.csr_col.p,
                               I.diag.p, I.lu.p);
            break;
        case 4:
            hipLaunchKernelGGL((k_ilu_factor_rows<4>), groups, KB, 0, c.stream, rows, cnt, c.rowptr.p, c.csr_col.p,
                               I.diag.p, I.lu.p);
            break;
        default:
            hipLaunchKernelGGL(k_ilu_factor, (cnt + KB - 1) / KB, KB, 0, c.stream, rows, cnt, c.rowptr.p,
                               c.csr_col.p, I.diag.p, I.lu.p);
        }
    }
    MHFEM_HIP(hipGetLastError());
}

// ParILU factors of the current matrix into k.ilu.lu (CSR order)
constexpr int PARILU_SWEEPS = 5;
static void parilu_factor(Ctx &c, Krylov &k)
{
    Ilu0 &I = k.ilu;
    const int n = k.n;
    const unsigned gr = (unsigned)((n + KB - 1) / KB);
    hipLaunchKernelGGL((k_gather<double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, c.sell_val.p, c.csr_to_sell.p,
                       I.a0.p);
    hipLaunchKernelGGL(k_parilu_init, gr, KB, 0, c.stream, n, c.rowptr.p, c.csr_col.p, I.diag.p, I.a0.p, I.lu.p);
    double *src = I.lu.p, *dst = I.swp.p;
    for (int s = 0; s < PARILU_SWEEPS; s++) {
        hipLaunchKernelGGL(k_parilu_sweep, gr, KB, 0, c.stream, n, c.rowptr.p, c.csr_col.p, I.diag.p, I.a0.p, src,
                           dst);
        std::swap(src, dst);
    }
    if (src != I.lu.p)
        MHFEM_HIP(hipMemcpyAsync(I.lu.p, src, (size_t)c.nnz * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    MHFEM_HIP(hipGetLastError());
}

static void precond_update(Ctx &c, Krylov &k)
{
    const int n = k.n;
    const unsigned g = grid_for(n);
    // the diagonal is 1 by construction: the Jacobi scaling is 1 / 1
    hipLaunchKernelGGL((k_fill_t<double>), g, KB, 0, c.stream, (long)n, k.dinv.p, 1.0);
    MHFEM_HIP(hipGetLastError());
    const bool mixed = c.cfg.precision != MHFEM_PRECISION_FP64;
    if (k.permuted64) {
        with_n(c.n, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            hipLaunchKernelGGL((k_convert_permuted<N, double>), grid_for(k.blocks32), KB, 0, c.stream, k.blocks32,
                               k.map32.p, c.sell_val.p, k.val64p.p);
            if constexpr (N > 1)
                hipLaunchKernelGGL((k_convert_permuted_diag<N, double>), grid_for(c.F), KB, 0, c.stream, c.F,
                                   k.perm.p, c.sell_val.p + c.diag_base, k.val64p.p + k.diag_base32);
        });
        MHFEM_HIP(hipGetLastError());
    }
    if (mixed && k.permuted) {
        with_n(c.n, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            hipLaunchKernelGGL((k_convert_permuted<N, float>), grid_for(k.blocks32), KB, 0, c.stream, k.blocks32,
                               k.map32.p, c.sell_val.p, k.val32.p);
            if constexpr (N > 1)
                hipLaunchKernelGGL((k_convert_permuted_diag<N, float>), grid_for(c.F), KB, 0, c.stream, c.F, k.perm.p,
                                   c.sell_val.p + c.diag_base, k.val32.p + k.diag_base32);
        });
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)n, k.dinv.p, k.dinv32.p);
        MHFEM_HIP(hipGetLastError());
    } else if (mixed) {
        hipLaunchKernelGGL((k_convert<float, double>), grid_for(c.sell_size), KB, 0, c.stream, c.sell_size,
                           c.sell_val.p, k.val32.p);
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)n, k.dinv.p, k.dinv32.p);
        MHFEM_HIP(hipGetLastError());
    }
    if (c.cfg.preconditioner == MHFEM_PREC_AMG) amg_update(c, *k.amg);
    if (is_isai(c.cfg.preconditioner)) {
        // rebuilt under the reuse rule of the reference's Ginkgo branch
        Isai &S = k.isai;
        const bool reuse = S.rebuilds > 0 && S.last_iters <= std::min(2 * S.updated_iters, S.updated_iters + 5);
        S.rebuilt = !reuse;
        if (!reuse) {
            if (c.cfg.preconditioner == MHFEM_PREC_PARILU_ISAI) parilu_factor(c, k);
            else ilu_factor(c, k);
            const Ilu0 &I = k.ilu;
            const unsigned gr = (unsigned)((n + KB - 1) / KB);
            hipLaunchKernelGGL(k_isai<true>, gr, KB, 0, c.stream, n, S.lo.rp.p, S.lo.col.p, c.rowptr.p, c.csr_col.p,
                               I.diag.p, I.lu.p, S.lo.val.p);
            hipLaunchKernelGGL(k_isai<false>, gr, KB, 0, c.stream, n, S.up.rp.p, S.up.col.p, c.rowptr.p, c.csr_col.p,
                               I.diag.p, I.lu.p, S.up.val.p);
            MHFEM_HIP(hipGetLastError());
            if (S.rebuilds == 0)
                for (IsaiFactor *F : {&S.lo, &S.up}) F->lanes = isai_tune_lanes(c, *F, k.dinv.p, S.t.p);
            S.rebuilds++;
        }
    }
    if (c.cfg.preconditioner == MHFEM_PREC_ILUT) {
        // every step, as the reference's preconditioners (Solver.hpp:1237-1242)
        Ilu0 &I = k.ilu;
        hipLaunchKernelGGL((k_gather<double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, c.sell_val.p,
                           c.csr_to_sell.p, I.a0.p);
        MHFEM_HIP(hipGetLastError());
        I.h_a.resize(c.nnz);
        MHFEM_HIP(hipMemcpyAsync(I.h_a.data(), I.a0.p, c.nnz * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        std::vector<int> frp, fcol;
        std::vector<double> fval;
        ilut_host(n, c.h_rowptr.data(), c.h_col.data(), I.h_a.data(), c.cfg.ilut_p, c.cfg.ilut_threshold, frp, fcol,
                  fval);
        if (I.lu.n < fval.size()) I.lu.alloc(fval.size());
        if (I.fcol.n < fcol.size()) I.fcol.alloc(fcol.size());
        if (!I.frp.p) I.frp.alloc(n + 1);
        MHFEM_HIP(hipMemcpy(I.frp.p, frp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(I.fcol.p, fcol.data(), fcol.size() * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(I.lu.p, fval.data(), fval.size() * sizeof(double), hipMemcpyHostToDevice));
        build_levels(I, n, frp.data(), fcol.data());
        I.own = true;
    }
    if (c.cfg.preconditioner == MHFEM_PREC_ILU0) {
        Ilu0 &I = k.ilu;
        ilu_factor(c, k);
        if (mixed)
            hipLaunchKernelGGL((k_convert<float, double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, I.lu.p, I.lu32.p);
        // the intermediate vectors start as sentinels, the tickets and the
        // error flag at zero (single-launch solves only)
        if (I.sweep_state.p)
            hipLaunchKernelGGL((k_ilu_prepare<double>), grid_for(n), KB, 0, c.stream, (long)n, I.zmid.p,
                               I.sweep_state.p, I.sweep_state.p + 2);
        if (mixed && I.sweep_state.p)
            hipLaunchKernelGGL((k_ilu_prepare<float>), grid_for(n), KB, 0, c.stream, (long)n, I.zmid32.p,
                               I.sweep_state.p, I.sweep_state.p + 2);
        MHFEM_HIP(hipGetLastError());
    }
}

Amg *krylov_amg(const Krylov &k) { return k.amg; }

bool krylov_isai_info(const Krylov &k, long *nnz_lower, long *nnz_upper, int *rebuilds)
{
    if (!k.isai.lo.rp.p) return false;
    if (nnz_lower) *nnz_lower = k.isai.lo.nnz;
    if (nnz_upper) *nnz_upper = k.isai.up.nnz;
    if (rebuilds) *rebuilds = k.isai.rebuilds;
    return true;
}

void krylov_isai_export(Ctx &c, int upper, int32_t *rowptr, int32_t *col, double *val)
{
    const IsaiFactor &F = upper ? c.krylov->isai.up : c.krylov->isai.lo;
    const int n = c.krylov->n;
    auto get = [&](void *dst, const void *src, size_t bytes) {
        if (dst) MHFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    };
    get(rowptr, F.rp.p, (n + 1) * sizeof(int));
    get(col, F.col.p, F.nnz * sizeof(int));
    get(val, F.val.p, F.nnz * sizeof(double));
    MHFEM_HIP(hipStreamSynchronize(c.stream));
}

void launch_csr_gather(Ctx &c, double *dst)
{
    hipLaunchKernelGGL((k_gather<double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, c.sell_val.p, c.csr_to_sell.p,
                       dst);
    MHFEM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// FP64 helpers for callers outside the Krylov loop (learned initial guess,
// benchmarks): no convergence flag is consulted.
void spmv64(Ctx &c, const double *x, double *y)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    S.spmv_plain<double>(x, y);
}

// AH = A H over the first `slots` slots of every row (rounded up to a
// multiple of 4)
void spmm_hist(Ctx &c, const double *H, double *AH, int slots)
{
    if (slots < 1) return;
    if (slots > HIST_SLOTS) throw Error{MHFEM_E_ARGUMENT, "spmm_hist: too many slots"};
    Krylov &k = *c.krylov;
    Solve S(c, k);
    SellView<double> A = S.matrix<double>();
    const int g4 = (slots + 3) / 4;
    const unsigned grid = (unsigned)(((int64_t)c.F * g4 + KB - 1) / KB);
    with_n(c.n, [&](auto NC) {
        constexpr int NN = decltype(NC)::value;
        switch (g4) {
        case 1: hipLaunchKernelGGL((k_spmm_hist<NN, 1>), grid, KB, 0, c.stream, A, H, AH); break;
        case 2: hipLaunchKernelGGL((k_spmm_hist<NN, 2>), grid, KB, 0, c.stream, A, H, AH); break;
        case 3: hipLaunchKernelGGL((k_spmm_hist<NN, 3>), grid, KB, 0, c.stream, A, H, AH); break;
        default: hipLaunchKernelGGL((k_spmm_hist<NN, 4>), grid, KB, 0, c.stream, A, H, AH); break;
        }
    });
    MHFEM_HIP(hipGetLastError());
}

void residual64(Ctx &c, const double *x, const double *b, double *r)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    S.residual<double>(x, b, r);
}

void residual64_halt(Ctx &c, const double *x, const double *b, double *r, const double *sc)
{
    SellView<double> A{c.F, c.slice_ptr.p, c.slice_width.p, c.sell_col.p, c.sell_val.p, c.diag_base};
    with_n(c.n, [&](auto NC) {
        hipLaunchKernelGGL((k_residual_halt<decltype(NC)::value>), grid_for(c.F), KB, 0, c.stream, A, x, b, r, sc);
    });
    MHFEM_HIP(hipGetLastError());
}

double dot64(Ctx &c, const double *x, const double *y)
{
    Krylov &k = *c.krylov;
    if (k.nseg != 1) throw Error{MHFEM_E_STATE, "dot64: single-segment systems only"};
    Solve S(c, k);
    S.dots<double, 1>(DotArgs<double, 1>{{x}, {y}}, FIN_NONE, M_ALL);
    S.poll();
    return k.h_sc[S_D0];
}

// Average duration of the solver's SpMV launch (y = A x on the current
// matrix), measured with HIP events on the context's stream.
void krylov_bench_spmv(Ctx &c, int precision, int reps, double *avg_ms)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    const unsigned g = grid_for(k.n);
    hipEvent_t e0 = c.ev[12], e1 = c.ev[13];
    if (precision != MHFEM_PRECISION_FP64) {
        if (!k.val32.p) throw Error{MHFEM_E_STATE, "FP32 matrix not configured"};
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)k.n, c.arr[MHFEM_Z_iF].p, k.x32.p);
    }
    for (int rep = -2; rep < reps; rep++) {
        if (rep == 0) MHFEM_HIP(hipEventRecord(e0, c.stream));
        if (precision != MHFEM_PRECISION_FP64)
            S.spmv_plain<float>(k.x32.p, k.tmp32.p);
        else
            S.spmv_plain<double>(c.arr[MHFEM_Z_iF].p, k.tmp.p);
    }
    MHFEM_HIP(hipGetLastError());
    MHFEM_HIP(hipEventRecord(e1, c.stream));
    MHFEM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MHFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = reps > 0 ? ms / reps : 0.0;
}

// Autotuner: rows per work-group of the segmented iteration kernels.  For
// each candidate it times a few real BiCGStab iterations -- both products,
// the vector updates and the reductions, in the precision the solve iterates
// in -- on the assembled system, from a scratch iterate (the solution is not
// touched; every solve starts its recurrence afresh), and keeps the fastest.
// A batched solve ends with its largest instance iterating alone, so the
// candidates are limited to those that still give that segment about two
// work-groups per CU (2 x 256).
int krylov_autotune(Ctx &c, double *best_ms)
{
    Krylov &k = *c.krylov;
    const int cands[] = {256, 512, 1024, 2048, 4096, 8192};
    hipEvent_t e0 = c.ev[12], e1 = c.ev[13];
    int best = k.rows_per_block;
    float bestms = 1e30f;
    const bool fp32 = c.cfg.precision != MHFEM_PRECISION_FP64 && k.val32.p;
    const int iters = c.cfg.preconditioner == MHFEM_PREC_ILU0 ? 2 : 8;
    int largest = 0;
    for (int sg = 0; sg < k.nseg; sg++) largest = std::max(largest, c.h_seg_rows[sg + 1] - c.h_seg_rows[sg]);
    const int limit = std::max(cands[0], largest / 512);
    const unsigned g = grid_for(k.n);
    precond_update(c, k);
    DBuf<double> xs;  // scratch FP64 iterate
    if (fp32) {
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)k.n, c.arr[MHFEM_RHS].p, k.rhs32.p);
    } else {
        xs.alloc(k.n);
    }
    MHFEM_HIP(hipGetLastError());
    for (int rpb : cands) {
        if (rpb > limit && rpb != cands[0]) continue;
        build_blocks(c, k, rpb);
        Solve S(c, k);
        auto run = [&](auto zero) {
            using T = decltype(zero);
            T *x;
            const T *b;
            Solve::Vecs v;
            if constexpr (sizeof(T) == 4) {
                x = k.x32.p;
                b = k.rhs32.p;
                v = Solve::Vecs{k.r32.p, k.r_ast32.p, k.p32.p, k.Ap32.p, k.s32.p, k.As32.p, k.tmp32.p, k.b32.p,
                                k.dinv32.p};
                hipLaunchKernelGGL((k_fill_t<float>), g, KB, 0, c.stream, (long)k.n, k.x32.p, 0.0f);
            } else {
                x = xs.p;
                b = c.arr[MHFEM_RHS].p;
                v = Solve::Vecs{k.r.p, k.r_ast.p, k.p.p, k.Ap.p, k.s.p, k.As.p, k.tmp.p, k.tmp2.p, k.dinv.p};
                MHFEM_HIP(hipMemcpyAsync(xs.p, c.arr[MHFEM_Z_iF].p, (size_t)k.n * sizeof(double),
                                         hipMemcpyDeviceToDevice, c.stream));
            }
            MHFEM_HIP(hipGetLastError());
            // tolerance 0: no segment stops during the measurement
            S.bicg_start<T>(b, x, v, 0.0, 1 << 30);
            S.bicg_iterations<T>(x, v, 2);
            MHFEM_HIP(hipEventRecord(e0, c.stream));
            S.bicg_iterations<T>(x, v, iters);
            MHFEM_HIP(hipEventRecord(e1, c.stream));
        };
        if (fp32) run(0.0f);
        else run(0.0);
        MHFEM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MHFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= iters;
        if (ms < bestms) {
            bestms = ms;
            best = rpb;
        }
    }
    build_blocks(c, k, best);
    drop_graphs(k);
    if (best_ms) *best_ms = bestms;
    return best;
}

void krylov_solve(Ctx &c, const double *b, double *x, mhfem_solve_stats *st)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    const int n = k.n;
    const unsigned g = grid_for(n);
    hipEvent_t e0 = c.ev[3], e1 = c.ev[4], e2 = c.ev[5];
    if (c.timing) MHFEM_HIP(hipEventRecord(e0, c.stream));
    precond_update(c, k);
    if (c.timing) MHFEM_HIP(hipEventRecord(e1, c.stream));
    const double tol = c.cfg.convergence_residue;
    const int maxit = c.cfg.max_iterations;
    mhfem_solve_stats out{};
    if (c.guess && c.guess_on_solve) guess_before_solve(c, *c.guess, b, x);
    if (k.nseg > 1 && (c.cfg.solver != MHFEM_SOLVER_BICGSTAB || c.cfg.precision != MHFEM_PRECISION_FP64))
        throw Error{MHFEM_E_ARGUMENT, "batched systems are solved with FP64 BiCGStab"};
    if (c.cfg.precision == MHFEM_PRECISION_FP64) {
        if (c.cfg.solver != MHFEM_SOLVER_BICGSTAB) {
            double res = 0;
            switch (c.cfg.solver) {
            case MHFEM_SOLVER_GMRES: out.iterations = S.gmres(b, x, tol, maxit, c.cfg.gmres_restart, &res); break;
            case MHFEM_SOLVER_CG: out.iterations = S.cg(b, x, tol, maxit, &res); break;
            case MHFEM_SOLVER_TFQMR: out.iterations = S.tfqmr(b, x, tol, maxit, &res); break;
            case MHFEM_SOLVER_BICGSTABL: out.iterations = S.bicgstabl(b, x, tol, maxit, c.cfg.bicgstab_ell, &res); break;
            case MHFEM_SOLVER_IDRS: out.iterations = S.idrs(b, x, tol, maxit, c.cfg.idr_s, &res); break;
            case MHFEM_SOLVER_SOR: out.iterations = S.sor(b, x, tol, maxit, c.cfg.sor_omega, &res); break;
            default: out.iterations = S.richardson(b, x, tol, maxit, &res); break;
            }
            out.residue = res;
            out.converged = res < tol;
        } else {
            Solve::Vecs v{k.r.p, k.r_ast.p, k.p.p, k.Ap.p, k.s.p, k.As.p, k.tmp.p, k.tmp2.p, k.dinv.p};
            if (k.permuted64) {
                with_n(c.n, [&](auto NC) {
                    constexpr int N = decltype(NC)::value;
                    hipLaunchKernelGGL((k_permute_faces<N, true>), grid_for(c.F), KB, 0, c.stream, c.F, k.perm.p, b,
                                       k.bp.p);
                    hipLaunchKernelGGL((k_permute_faces<N, true>), grid_for(c.F), KB, 0, c.stream, c.F, k.perm.p,
                                       (const double *)x, k.xp.p);
                });
                MHFEM_HIP(hipGetLastError());
                S.perm64 = true;
                out.iterations = S.bicgstab<double>(k.bp.p, k.xp.p, v, tol, maxit);
                S.perm64 = false;
                with_n(c.n, [&](auto NC) {
                    constexpr int N = decltype(NC)::value;
                    hipLaunchKernelGGL((k_permute_faces<N, false>), grid_for(c.F), KB, 0, c.stream, c.F, k.perm.p,
                                       (const double *)k.xp.p, x);
                });
                MHFEM_HIP(hipGetLastError());
            } else {
                out.iterations = S.bicgstab<double>(b, x, v, tol, maxit);
            }
            out.residue = 0;
            out.converged = 1;
            for (int s = 0; s < k.nseg; s++) {
                out.residue = std::max(out.residue, S.seg_sc(s)[S_RESID]);
                out.converged &= S.seg_sc(s)[S_RESID] < tol;
            }
        }
        out.outer_iterations = 1;
    } else {
        // FP32 inner BiCGStab inside FP64 iterative refinement.  Outer test:
        // ||M^{-1}(b - A x)|| / ||M^{-1} b|| < tol, with M the FP32 preconditioner
        // applied to the FP64 residual rounded to FP32.
        float *rhs32 = k.rhs32.p, *z = k.b32.p;
        Solve::Vecs v{k.r32.p, k.r_ast32.p, k.p32.p, k.Ap32.p, k.s32.p, k.As32.p, k.tmp32.p, k.b32.p, k.dinv32.p};
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)n, b, rhs32);
        S.prec<float>(rhs32, z, k.dinv32.p, M_ALL);
        S.dots<float, 1>(DotArgs<float, 1>{{z}, {z}}, FIN_NONE, M_ALL);
        S.poll();
        double b_norm = std::sqrt(k.h_sc[S_D0]);
        if (b_norm == 0.0) b_norm = 1.0;
        const bool reliable = c.cfg.precision == MHFEM_PRECISION_FP32_RELIABLE;
        // After a solve whose FP32 iteration stalled, the next solves of the
        // trajectory (nearby systems) go straight to FP64 for a while.
        constexpr int cooldown = 8;
        bool fallback = k.fp64_cooldown > 0;
        const bool skip_fp32 = fallback;
        if (skip_fp32) k.fp64_cooldown--;
        const double inner_tol = c.cfg.inner_residue > 0 ? c.cfg.inner_residue : (reliable ? 1e-2 : 1e-4);
        int total = 0, outer = 0;
        double res = 0;
        for (; reliable && !skip_fp32;) {
            // One FP32 recurrence with reliable updates (Sleijpen & van der
            // Vorst 1996): each time its residual has dropped by inner_tol,
            // the FP32 correction goes into the FP64 solution, the FP64
            // residual replaces the recursive one and the recurrence goes on
            // with its search direction and shadow residual.
            float *r = k.r32.p;
            S.residual_mixed(x, b, rhs32);
            S.prec<float>(rhs32, z, k.dinv32.p, M_ALL);
            S.dots<float, 1>(DotArgs<float, 1>{{z}, {z}}, FIN_NONE, M_ALL);
            S.poll();
            res = std::sqrt(k.h_sc[S_D0]) / b_norm;
            if (!(res >= tol)) break;
            hipLaunchKernelGGL((k_fill_t<float>), g, KB, 0, c.stream, (long)n, k.x32.p, 0.0f);
            S.bicg_start<float>(rhs32, k.x32.p, v, std::max(inner_tol, tol / res), maxit);
            hipLaunchKernelGGL(k_set_slots, 1, 64, 0, c.stream, k.sc.p, tol * b_norm, inner_tol);
            // each round between two replacements gets at most `round_cap`
            // iterations, so that a recurrence that cannot reach the next
            // replacement still reaches the stall test below
            const int round_cap = 300;
            hipLaunchKernelGGL(k_set_slot, 1, 64, 0, c.stream, k.sc.p, (int)S_MAXIT, (double)std::min(maxit, round_cap));
            MHFEM_HIP(hipGetLastError());
            const int every = std::max(1, c.cfg.check_every);
            int stalls = 0;
            for (;;) {
                total = S.bicg_drive<float>(rhs32, k.x32.p, v,
                                            [&]() { S.bicg_iterations<float>(k.x32.p, v, every); });
                if (!std::isfinite(k.h_sc[S_RESID])) {  // a broken FP32 recurrence is not flushed
                    fallback = true;
                    break;
                }
                S.add_correction(x, k.x32.p);
                // the merged iteration has advanced p with the recursive
                // residual: keep p - r for the continued direction
                const bool merged = S.merged<float>();
                if (merged)
                    hipLaunchKernelGGL((k_bicg_dir<float>), k.nblk, KB, 0, c.stream, S.seg(M_ALL), (const float *)k.p32.p,
                                       (const float *)r, k.s32.p);
                S.residual_mixed(x, b, rhs32);
                MHFEM_HIP(hipGetLastError());
                S.prec<float>(rhs32, r, k.dinv32.p, M_ALL);
                // merged: D0 = (r, r) (read below), (r, r) again in partial
                // row 1 for the next merged alpha, D2 = (r, r*)
                if (merged)
                    S.dots<float, 3>(DotArgs<float, 3>{{r, r, r}, {r, r, k.r_ast32.p}}, FIN_REPLACE_MERGED, M_ALL);
                else
                    S.dots<float, 2>(DotArgs<float, 2>{{r, r}, {r, k.r_ast32.p}}, FIN_REPLACE, M_ALL);
                S.poll();
                outer++;
                const double prev = res;
                res = std::sqrt(k.h_sc[S_D0]) / b_norm;
                if (!(res >= tol) || total >= maxit) break;
                // FP32 iterations only refine while kappa * eps_fp32 << 1: two
                // replacements in a row that do not halve the FP64 residual
                // hand the rest of the solve to FP64 BiCGStab
                stalls = res > 0.5 * prev ? stalls + 1 : 0;
                if (stalls >= 2) {
                    fallback = true;
                    break;
                }
                hipLaunchKernelGGL(k_set_slot, 1, 64, 0, c.stream, k.sc.p, (int)S_MAXIT,
                                   (double)std::min(maxit, total + round_cap));
                if (merged)
                    hipLaunchKernelGGL((k_bicg_pfix<float>), S.sg.nact, KB, 0, c.stream, S.sg, (const float *)r,
                                       k.p32.p, (const float *)k.s32.p);
                else
                    hipLaunchKernelGGL((k_bicg_p<float>), S.sg.nact, KB, 0, c.stream, S.sg, (const float *)r, k.p32.p,
                                       (const float *)k.Ap32.p);
                MHFEM_HIP(hipGetLastError());
            }
            break;
        }
        double res_before = 0.0;
        for (; !reliable && !skip_fp32;) {
            S.residual_mixed(x, b, rhs32);
            S.prec<float>(rhs32, z, k.dinv32.p, M_ALL);
            S.dots<float, 1>(DotArgs<float, 1>{{z}, {z}}, FIN_NONE, M_ALL);
            S.poll();
            res = std::sqrt(k.h_sc[S_D0]) / b_norm;
            if (!(res >= tol) || total >= maxit) break;
            if (outer >= 2 && res > 0.5 * res_before) {  // refinement no longer converging: FP64 takes over
                fallback = true;
                break;
            }
            res_before = res;
            hipLaunchKernelGGL((k_fill_t<float>), g, KB, 0, c.stream, (long)n, k.x32.p, 0.0f);
            MHFEM_HIP(hipGetLastError());
            total += S.bicgstab<float>(rhs32, k.x32.p, v, inner_tol, std::max(1, maxit - total));
            if (!std::isfinite(k.h_sc[S_RESID])) {
                fallback = true;
                break;
            }
            S.add_correction(x, k.x32.p);
            MHFEM_HIP(hipGetLastError());
            outer++;
        }
        if (fallback && !skip_fp32) k.fp64_cooldown = cooldown;
        if (fallback && total < maxit) {
            // FP64 BiCGStab from the current iterate; its stopping test is the
            // plain method's
            Solve::Vecs v64{k.r.p, k.r_ast.p, k.p.p, k.Ap.p, k.s.p, k.As.p, k.tmp.p, k.tmp2.p, k.dinv.p};
            const int it64 = S.bicgstab<double>(b, x, v64, tol, std::max(1, maxit - total));
            total += it64;
            out.fallback_iterations = it64;
            res = S.seg_sc(0)[S_RESID];
        }
        out.iterations = total;
        out.outer_iterations = outer;
        out.residue = res;
        out.converged = res < tol;
    }
    if (c.cfg.preconditioner == MHFEM_PREC_ILU0 && k.ilu.sweep_state.p) {
        int err = 0;
        MHFEM_HIP(hipMemcpyAsync(&err, k.ilu.sweep_state.p + 2, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        if (err)
            throw Error{MHFEM_E_STATE, "ILU(0) triangular solve: a wait for a dependency exceeded its limit (the "
                                       "values of this solve are not valid)"};
    }
    if (c.guess && c.guess_on_solve) guess_after_solve(c, *c.guess, x);
    if (k.amg) amg_after_solve(*k.amg, out.iterations);
    if (is_isai(c.cfg.preconditioner)) {
        k.isai.last_iters = out.iterations;
        if (k.isai.rebuilt || k.isai.updated_iters == 0) k.isai.updated_iters = out.iterations;
    }
    if (c.timing) {
        MHFEM_HIP(hipEventRecord(e2, c.stream));
        MHFEM_HIP(hipEventSynchronize(e2));
        float t1 = 0, t2 = 0;
        MHFEM_HIP(hipEventElapsedTime(&t1, e0, e1));
        MHFEM_HIP(hipEventElapsedTime(&t2, e1, e2));
        c.timers.preconditioner += t1;
        c.timers.linear_solver += t2;
    }
    *st = out;
}

// per-segment results of the last BiCGStab solve
void krylov_segment_stats(Ctx &c, int *iterations, double *residues, int count)
{
    Krylov &k = *c.krylov;
    for (int s = 0; s < std::min(count, k.nseg); s++) {
        if (iterations) iterations[s] = (int)k.h_sc[(long)s * NSLOTS + S_IT];
        if (residues) residues[s] = k.h_sc[(long)s * NSLOTS + S_RESID];
    }
}

// ---------------------------------------------------------------------------
// One time step as captured HIP graphs (FP64 BiCGStab):
//   first = preIterate + assembly + preconditioner update + BiCGStab start
//           + `check_every` iterations,
//   more  = `check_every` further iterations (replayed until every segment
//           has converged),
//   post  = postIterate.
// Scalars are initialised by a kernel and converged segments turn the queued
// work into no-ops, so replaying the same graphs is exact.
void krylov_graph_step(Ctx &c, double time, double tau, mhfem_solve_stats *st)
{
    Krylov &k = *c.krylov;
    if (c.cfg.solver != MHFEM_SOLVER_BICGSTAB || c.cfg.precision != MHFEM_PRECISION_FP64)
        throw Error{MHFEM_E_ARGUMENT, "graph steps use FP64 BiCGStab"};
    if (c.cfg.preconditioner == MHFEM_PREC_AMG || is_isai(c.cfg.preconditioner) ||
        c.cfg.preconditioner == MHFEM_PREC_ILUT)
        throw Error{MHFEM_E_ARGUMENT, "graph steps do not take the AMG, ILU-ISAI or ILUT preconditioners (their set-up "
                                      "decisions run on the host); use the four calls"};
    const int every = std::max(1, c.cfg.check_every);
    // The graphs hold the kernel arguments of their capture: tau, the batch
    // size and every device pointer and uniform mask of the scheme.  Any
    // change (a coefficient set uniform or full, a reallocated array, a
    // boundary model switched on) re-captures them.  SchemeView is built
    // zero-initialised, so its padding compares equal.
    const SchemeView view = c.view();
    if (k.g_first && (k.g_tau != tau || k.g_iters != every || std::memcmp(&k.g_view, &view, sizeof view) != 0))
        drop_graphs(k);
    Solve S(c, k);
    Solve::Vecs v{k.r.p, k.r_ast.p, k.p.p, k.Ap.p, k.s.p, k.As.p, k.tmp.p, k.tmp2.p, k.dinv.p};
    const double *b = c.arr[MHFEM_RHS].p;
    double *x = c.arr[MHFEM_Z_iF].p;
    auto capture = [&](auto &&body, hipGraphExec_t *out) {
        const bool timing = c.timing;
        c.timing = false;  // no event records inside a captured graph
        hipGraph_t graph;
        MHFEM_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        body();
        MHFEM_HIP(hipStreamEndCapture(c.stream, &graph));
        MHFEM_HIP(hipGraphInstantiate(out, graph, nullptr, nullptr, 0));
        MHFEM_HIP(hipGraphDestroy(graph));
        c.timing = timing;
    };
    if (!k.g_first) {
        capture([&]() {
            launch_pre_iterate(c, time, tau);
            launch_assemble(c, time, tau);
            precond_update(c, k);
            S.bicg_start<double>(b, x, v, c.cfg.convergence_residue, c.cfg.max_iterations);
            S.bicg_iterations<double>(x, v, every);
        }, &k.g_first);
        capture([&]() { S.bicg_iterations<double>(x, v, every); }, &k.g_more);
        capture([&]() { launch_post_iterate(c); }, &k.g_post);
        k.g_tau = tau;
        k.g_iters = every;
        k.g_view = view;
        k.g_stream = c.stream;
    }
    // the boundary model's time is read from device memory by the captured
    // assembly, so each replay sees its own step's time + tau
    launch_set_clock(c, time + tau);
    MHFEM_HIP(hipGraphLaunch(k.g_first, c.stream));
    // once the launches are narrowed to the instances still iterating, the
    // same iterations go eagerly over the narrowed grids (the kernels and
    // arguments of the capture but for the grid, so the same results)
    const int it = S.bicg_drive<double>(b, x, v, [&]() {
        if (S.sg.active) S.bicg_iterations<double>(x, v, every);
        else MHFEM_HIP(hipGraphLaunch(k.g_more, c.stream));
    });
    MHFEM_HIP(hipGraphLaunch(k.g_post, c.stream));
    mhfem_solve_stats out{};
    out.iterations = it;
    out.outer_iterations = 1;
    out.converged = 1;
    for (int s = 0; s < k.nseg; s++) {
        out.residue = std::max(out.residue, S.seg_sc(s)[S_RESID]);
        out.converged &= S.seg_sc(s)[S_RESID] < c.cfg.convergence_residue;
    }
    *st = out;
}

}  // namespace mhfem
