// krylov.hip -- linear solvers for the hybridised MHFEM system on gfx950.
//
// * SELL-64 SpMV: one lane per row, a wavefront covers one 64-row slice, so
//   every matrix load of a wavefront is one contiguous 512-byte segment.
// * BiCGStab (left preconditioned, van der Vorst 1992) as five fused kernels
//   per iteration; dot products reduce wavefront -> work-group -> grid
//   (deterministic order) and the scalar recurrences
//   (alpha, omega, beta, residue, convergence) are evaluated on the device, so
//   the host only polls a status word every `check_every` iterations.
//   Partial sums are finished by a one-work-group kernel on the same stream.  Kernels
//   queued after convergence see the DONE flag and return immediately.
// * GMRES(m), left preconditioned, classical Gram-Schmidt with one
//   re-orthogonalisation pass; Givens rotations on the host.
// * Jacobi (diagonal) and ILU(0) (Saad Alg. 10.4, exact, level-scheduled:
//   rows of one dependency level are processed by one launch).
// * FP32 inner BiCGStab inside FP64 iterative refinement.
//
// Solver call sites in the reference: Solver.hpp:1230-1265 (TNL LinearSolver +
// Preconditioner chosen by name, configSetup.h:31-62).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace mhfem {

constexpr int KB = 256;
constexpr int MAXG = 2048;

// device scalar slots
enum Slot {
    S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_BETA, S_RAST_SQ, S_BNORM, S_RESID, S_TOL, S_IT,
    S_DONE, S_RESTART, S_MAXIT, S_D0, S_D1, S_D2, S_D3, NSLOTS = 32
};
enum Fin { FIN_NONE = 0, FIN_ALPHA, FIN_OMEGA, FIN_UPDATE, FIN_BNORM, FIN_INIT };

struct Red {
    double *partials;  // one row of per-work-group partial sums per dot product
    double *sc;        // device scalar slots
};

__device__ __forceinline__ bool halted(const double *sc) { return sc[S_DONE] != 0.0 || sc[S_RESTART] != 0.0; }

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
    case FIN_ALPHA: sc[S_ALPHA] = sc[S_RHO] / sc[S_D0]; break;
    case FIN_OMEGA: sc[S_OMEGA] = sc[S_D1] > 0 ? sc[S_D0] / sc[S_D1] : 0.0; break;
    case FIN_UPDATE: {
        sc[S_RHO_OLD] = sc[S_RHO];
        sc[S_RHO] = sc[S_D0];
        sc[S_IT] += 1;
        if (fabs(sc[S_RHO]) < eps2 * sc[S_RAST_SQ]) {
            sc[S_RESTART] = 1;  // the host recomputes the residual (restart)
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
        sc[S_RHO] = sc[S_D0];
        sc[S_RAST_SQ] = sc[S_D0];
        const double res = sqrt(sc[S_D0]) / sc[S_BNORM];
        sc[S_RESID] = res;
        sc[S_RESTART] = 0;
        sc[S_DONE] = (!(res >= sc[S_TOL]) || sc[S_IT] >= sc[S_MAXIT]) ? 1 : 0;
        break;
    }
    default: break;
    }
}

// Reduce ND per-thread values to one partial sum per work-group
// (wavefront shuffles, then LDS).  The partials are summed by k_finalize,
// launched next on the same stream: a kernel boundary orders the writes, so
// no work-group issues a device-scope fence (which on gfx950 writes back the
// L2 of its XCD and, once per work-group, costs more than the kernel's own
// streaming).
template <int ND>
__device__ __forceinline__ void grid_reduce(double (&v)[ND], Red red, int fin)
{
    (void)fin;
    __shared__ double sm[ND][KB / 64];
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
            for (int q = 1; q < KB / 64; q++) s += sm[d][q];
            red.partials[(long)d * gridDim.x + blockIdx.x] = s;
        }
    }
}

// Sum the per-work-group partials of the previous kernel into sc[S_D0 + d]
// and evaluate the scalar recurrence `fin`.  One work-group.
template <int ND>
__global__ void __launch_bounds__(KB) k_finalize(int G, Red red, int fin, int check_halt)
{
    if (check_halt && halted(red.sc)) return;
    __shared__ double sm[ND][KB / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        double a = 0;
        for (int b = tid; b < G; b += KB) a += red.partials[(long)d * G + b];
        a = wave_sum(a);
        if (lane == 0) sm[d][wid] = a;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
            double s = sm[d][0];
#pragma unroll
            for (int q = 1; q < KB / 64; q++) s += sm[d][q];
            red.sc[S_D0 + d] = s;
        }
        finalize(red.sc, fin);
    }
}

// ---------------------------------------------------------------------------
// SpMV: y = D (A x) with D = dinv (Jacobi) or identity; optional dot (y, z).
template <typename T>
__device__ __forceinline__ T sell_row(const SellView<T> &A, const T *__restrict__ x, int row)
{
    const int64_t base = A.slice_ptr[row >> 6] + (row & 63);
    const int w = A.slice_width[row >> 6];
    T sum = 0;
    for (int k = 0; k < w; k++) {
        const int64_t q = base + (int64_t)k * 64;
        sum += A.val[q] * x[A.col[q]];
    }
    return sum;
}

template <typename T, bool DIAG, int ND>
__global__ void __launch_bounds__(KB) k_spmv(SellView<T> A, const T *__restrict__ x, const T *__restrict__ dinv,
                                             T *__restrict__ y, const T *__restrict__ z, Red red, int fin)
{
    // ND = 0: no dot; 1: (y, z); 2: (y, z) and (y, y)
    if (halted(red.sc)) return;
    double acc[ND > 0 ? ND : 1] = {};
    for (int row = blockIdx.x * KB + threadIdx.x; row < A.rows; row += gridDim.x * KB) {
        T s = sell_row(A, x, row);
        if (DIAG) s *= dinv[row];
        y[row] = s;
        if constexpr (ND >= 1) acc[0] += (double)s * (double)z[row];
        if constexpr (ND >= 2) acc[1] += (double)s * (double)s;
    }
    if constexpr (ND > 0) grid_reduce<ND>(acc, red, fin);
}

// y = b - A x  (residual, then preconditioned by the caller)
template <typename T>
__global__ void __launch_bounds__(KB) k_residual(SellView<T> A, const T *__restrict__ x, const T *__restrict__ b,
                                                 T *__restrict__ y)
{
    for (int row = blockIdx.x * KB + threadIdx.x; row < A.rows; row += gridDim.x * KB)
        y[row] = b[row] - sell_row(A, x, row);
}

// FP64 residual of the refinement loop, stored in FP32: r32 = (float)(b - A x)
__global__ void __launch_bounds__(KB) k_residual_mixed(SellView<double> A, const double *__restrict__ x,
                                                       const double *__restrict__ b, float *__restrict__ r32)
{
    for (int row = blockIdx.x * KB + threadIdx.x; row < A.rows; row += gridDim.x * KB)
        r32[row] = (float)(b[row] - sell_row(A, x, row));
}

template <typename T, int ND>
struct DotArgs {
    const T *x[ND];
    const T *y[ND];
};

template <typename T, int ND>
__global__ void __launch_bounds__(KB) k_dots(int n, DotArgs<T, ND> a, Red red, int fin, int check_halt)
{
    if (check_halt && halted(red.sc)) return;
    double acc[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) acc[d] = 0;
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB)
#pragma unroll
        for (int d = 0; d < ND; d++) acc[d] += (double)a.x[d][i] * (double)a.y[d][i];
    grid_reduce<ND>(acc, red, fin);
}

// s = r - alpha Ap
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_s(int n, const T *__restrict__ r, const T *__restrict__ Ap,
                                               T *__restrict__ s, const double *sc)
{
    if (halted(sc)) return;
    const T alpha = (T)sc[S_ALPHA];
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) s[i] = r[i] - alpha * Ap[i];
}

// x += alpha p + omega s; r = s - omega As; dots (r, r_ast), (r, r)
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_xr(int n, T *__restrict__ x, const T *__restrict__ p,
                                                const T *__restrict__ s, const T *__restrict__ As,
                                                T *__restrict__ r, const T *__restrict__ r_ast, Red red)
{
    if (halted(red.sc)) return;
    const T alpha = (T)red.sc[S_ALPHA], omega = (T)red.sc[S_OMEGA];
    double acc[2] = {0, 0};
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        const T si = s[i];
        x[i] += alpha * p[i] + omega * si;
        const T ri = si - omega * As[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)r_ast[i];
        acc[1] += (double)ri * (double)ri;
    }
    grid_reduce<2>(acc, red, FIN_UPDATE);
}

// p = r + beta (p - omega Ap)
template <typename T>
__global__ void __launch_bounds__(KB) k_bicg_p(int n, const T *__restrict__ r, T *__restrict__ p,
                                               const T *__restrict__ Ap, const double *sc)
{
    if (halted(sc)) return;
    const T beta = (T)sc[S_BETA], omega = (T)sc[S_OMEGA];
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) p[i] = r[i] + beta * (p[i] - omega * Ap[i]);
}

template <typename T>
__global__ void __launch_bounds__(KB) k_copy2(int n, const T *__restrict__ src, T *__restrict__ a, T *__restrict__ b)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        const T v = src[i];
        a[i] = v;
        if (b) b[i] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(KB) k_scale(int n, const T *__restrict__ x, const T *__restrict__ d, T *__restrict__ y)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) y[i] = d ? x[i] * d[i] : x[i];
}

template <typename TO, typename TI>
__global__ void __launch_bounds__(KB) k_convert(long n, const TI *__restrict__ x, TO *__restrict__ y)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) y[i] = (TO)x[i];
}

// x64 += (double) d32
__global__ void __launch_bounds__(KB) k_add_correction(int n, double *__restrict__ x, const float *__restrict__ d)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) x[i] += (double)d[i];
}

template <typename T>
__global__ void __launch_bounds__(KB) k_fill_t(long n, T *__restrict__ x, T v)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) x[i] = v;
}

// Jacobi: dinv[r] = 1 / A(r,r) (TNL Diagonal preconditioner)
template <typename T>
__global__ void __launch_bounds__(KB) k_diag_inv(int n, const T *__restrict__ val, const int64_t *__restrict__ dpos,
                                                 T *__restrict__ dinv)
{
    for (int i = blockIdx.x * KB + threadIdx.x; i < n; i += gridDim.x * KB) {
        const T d = val[dpos[i]];
        dinv[i] = d != (T)0 ? (T)1 / d : (T)1;
    }
}

// ---------------------------------------------------------------------------
// ILU(0), level-scheduled.  Factors are kept in CSR order (natural rows).
template <typename T>
__global__ void __launch_bounds__(KB) k_gather(long n, const T *__restrict__ src, const int64_t *__restrict__ map,
                                               T *__restrict__ dst)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB) dst[i] = src[map[i]];
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

struct Ilu0 {
    int n = 0;
    long nnz = 0;
    std::vector<int> h_lptr, h_uptr;  // level pointers
    DBuf<int> lrows, urows, diag;     // rows in level order, diagonal positions
    DBuf<double> lu;
    DBuf<float> lu32;
};

// ---------------------------------------------------------------------------
struct Krylov {
    int n = 0;
    DBuf<double> r, r_ast, p, Ap, s, As, tmp, tmp2, dinv;
    DBuf<int64_t> diag_pos;  // SELL position of the diagonal of each row
    DBuf<double> partials;
    DBuf<double> sc;
    double *h_sc = nullptr;  // pinned
    Ilu0 ilu;
    // FP32 inner solver
    DBuf<float> val32, dinv32, b32, x32, r32, r_ast32, p32, Ap32, s32, As32, tmp32, rhs32;
    // GMRES
    DBuf<double> V;
    int V_m = 0;
};

static inline unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + KB - 1) / KB, MAXG)); }

Krylov *krylov_create(Ctx &c)
{
    Krylov *k = new Krylov();
    const int n = c.rows;
    k->n = n;
    for (DBuf<double> *b : {&k->r, &k->r_ast, &k->p, &k->Ap, &k->s, &k->As, &k->tmp, &k->tmp2, &k->dinv})
        b->alloc(n);
    k->partials.alloc(4 * MAXG);
    k->sc.alloc(NSLOTS);
    MHFEM_HIP(hipHostMalloc((void **)&k->h_sc, NSLOTS * sizeof(double), hipHostMallocDefault));
    // diagonal positions in SELL storage
    std::vector<int64_t> dpos(n);
    for (int r = 0; r < n; r++) {
        int kk = -1;
        for (int q = c.h_rowptr[r]; q < c.h_rowptr[r + 1]; q++)
            if (c.h_col[q] == r) kk = q - c.h_rowptr[r];
        if (kk < 0) throw Error{MHFEM_E_STATE, "matrix row without a diagonal entry"};
        dpos[r] = sell_pos(c.h_slice_ptr.data(), r, kk);
    }
    k->diag_pos.alloc(n);
    MHFEM_HIP(hipMemcpy(k->diag_pos.p, dpos.data(), n * sizeof(int64_t), hipMemcpyHostToDevice));

    if (c.cfg.preconditioner == MHFEM_PREC_ILU0) {
        Ilu0 &I = k->ilu;
        I.n = n;
        I.nnz = c.nnz;
        std::vector<int> dg(n), lev(n), ulev(n);
        for (int r = 0; r < n; r++)
            for (int q = c.h_rowptr[r]; q < c.h_rowptr[r + 1]; q++)
                if (c.h_col[q] == r) dg[r] = q;
        int nl = 0, nu = 0;
        for (int r = 0; r < n; r++) {
            int l = 0;
            for (int q = c.h_rowptr[r]; q < dg[r]; q++) l = std::max(l, lev[c.h_col[q]] + 1);
            lev[r] = l;
            nl = std::max(nl, l + 1);
        }
        for (int r = n - 1; r >= 0; r--) {
            int l = 0;
            for (int q = dg[r] + 1; q < c.h_rowptr[r + 1]; q++) l = std::max(l, ulev[c.h_col[q]] + 1);
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
        I.lrows.alloc(n);
        I.urows.alloc(n);
        I.diag.alloc(n);
        MHFEM_HIP(hipMemcpy(I.lrows.p, lrows.data(), n * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(I.urows.p, urows.data(), n * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(I.diag.p, dg.data(), n * sizeof(int), hipMemcpyHostToDevice));
        I.lu.alloc(c.nnz);
        if (c.cfg.precision == MHFEM_PRECISION_FP32_REFINED) I.lu32.alloc(c.nnz);
    }
    if (c.cfg.precision == MHFEM_PRECISION_FP32_REFINED) {
        k->val32.alloc(c.sell_size);
        for (DBuf<float> *b : {&k->dinv32, &k->b32, &k->x32, &k->r32, &k->r_ast32, &k->p32, &k->Ap32, &k->s32,
                               &k->As32, &k->tmp32, &k->rhs32})
            b->alloc(n);
    }
    return k;
}

void krylov_destroy(Krylov *k)
{
    if (!k) return;
    if (k->h_sc) (void)hipHostFree(k->h_sc);
    delete k;
}

// ---------------------------------------------------------------------------
// host-side helpers
namespace {

struct Solve {
    Ctx &c;
    Krylov &k;
    int n;
    unsigned g;
    Red red;

    Solve(Ctx &c_, Krylov &k_) : c(c_), k(k_), n(k_.n), g(grid_for(k_.n))
    {
        red = Red{k.partials.p, k.sc.p};
    }

    template <typename T>
    SellView<T> matrix();

    void set_scalars(double tol, int maxit)
    {
        double h[NSLOTS] = {};
        h[S_TOL] = tol;
        h[S_MAXIT] = maxit;
        MHFEM_HIP(hipMemcpyAsync(k.sc.p, h, sizeof(h), hipMemcpyHostToDevice, c.stream));
    }
    void poll()
    {
        MHFEM_HIP(hipMemcpyAsync(k.h_sc, k.sc.p, NSLOTS * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    }
    void clear_flag(int slot)
    {
        const double z = 0;
        MHFEM_HIP(hipMemcpyAsync(k.sc.p + slot, &z, sizeof(double), hipMemcpyHostToDevice, c.stream));
    }

    template <typename T, int ND>
    void dots(DotArgs<T, ND> a, int fin, bool check_halt)
    {
        hipLaunchKernelGGL((k_dots<T, ND>), g, KB, 0, c.stream, n, a, red, fin, check_halt ? 1 : 0);
        finish<ND>(fin, check_halt);
    }
    // sum the partials of the reducing kernel just launched (grid g)
    template <int ND>
    void finish(int fin, bool check_halt)
    {
        hipLaunchKernelGGL((k_finalize<ND>), 1, KB, 0, c.stream, (int)g, red, fin, check_halt ? 1 : 0);
        MHFEM_HIP(hipGetLastError());
    }

    // ---- preconditioner apply y = M^{-1} x (T = double or float)
    template <typename T>
    void ilu_apply(const T *x, T *y, const double *sc = nullptr)
    {
        Ilu0 &I = k.ilu;
        const T *lu;
        if constexpr (sizeof(T) == 8) lu = I.lu.p;
        else lu = I.lu32.p;
        const int *rp = c.rowptr.p, *col = c.csr_col.p;
        for (size_t l = 0; l + 1 < I.h_lptr.size(); l++) {
            const int cnt = I.h_lptr[l + 1] - I.h_lptr[l];
            hipLaunchKernelGGL((k_ilu_lower<T>), (cnt + KB - 1) / KB, KB, 0, c.stream, I.lrows.p + I.h_lptr[l], cnt,
                               rp, col, I.diag.p, lu, x, y, sc);
        }
        for (size_t l = 0; l + 1 < I.h_uptr.size(); l++) {
            const int cnt = I.h_uptr[l + 1] - I.h_uptr[l];
            hipLaunchKernelGGL((k_ilu_upper<T>), (cnt + KB - 1) / KB, KB, 0, c.stream, I.urows.p + I.h_uptr[l], cnt,
                               rp, col, I.diag.p, lu, y, sc);
        }
        MHFEM_HIP(hipGetLastError());
    }
    // The assembled MHFEM rows are scaled so that every diagonal entry is
    // exactly 1 (LinearSystem.h:196-201, BoundaryConditions.hpp:200-203,
    // :375-377): the diagonal preconditioner is then the identity and its
    // scaling vector is not read.
    bool jacobi_scales() const { return c.cfg.preconditioner == MHFEM_PREC_DIAGONAL && !c.unit_diagonal; }

    template <typename T>
    void prec(const T *x, T *y, const T *dinv)
    {
        if (c.cfg.preconditioner == MHFEM_PREC_ILU0) ilu_apply<T>(x, y);
        else hipLaunchKernelGGL((k_scale<T>), g, KB, 0, c.stream, n, x, jacobi_scales() ? dinv : (const T *)nullptr, y);
        MHFEM_HIP(hipGetLastError());
    }

    // y = M^{-1} A x; ND fused dots: 1 -> (y, z), 2 -> (y, z), (y, y); then `fin`
    template <typename T, int ND>
    void pmatvec(const T *x, T *y, T *tmp, const T *dinv, const T *z, int fin)
    {
        SellView<T> A = matrix<T>();
        const int pc = c.cfg.preconditioner;
        if (pc == MHFEM_PREC_ILU0) {
            hipLaunchKernelGGL((k_spmv<T, false, 0>), g, KB, 0, c.stream, A, x, (const T *)nullptr, tmp,
                               (const T *)nullptr, red, 0);
            ilu_apply<T>(tmp, y, k.sc.p);
            if constexpr (ND == 1) dots<T, 1>(DotArgs<T, 1>{{y}, {z}}, fin, true);
            if constexpr (ND == 2) dots<T, 2>(DotArgs<T, 2>{{y, y}, {z, y}}, fin, true);
        } else if (jacobi_scales()) {
            hipLaunchKernelGGL((k_spmv<T, true, ND>), g, KB, 0, c.stream, A, x, dinv, y, z, red, fin);
            if constexpr (ND > 0) finish<ND>(fin, true);
        } else {
            hipLaunchKernelGGL((k_spmv<T, false, ND>), g, KB, 0, c.stream, A, x, dinv, y, z, red, fin);
            if constexpr (ND > 0) finish<ND>(fin, true);
        }
        MHFEM_HIP(hipGetLastError());
    }
    // r = M^{-1}(b - A x)
    template <typename T>
    void presidue(const T *x, const T *b, T *r, T *tmp, const T *dinv)
    {
        hipLaunchKernelGGL((k_residual<T>), g, KB, 0, c.stream, matrix<T>(), x, b, tmp);
        MHFEM_HIP(hipGetLastError());
        prec<T>(tmp, r, dinv);
    }

    struct Vecs {
        void *r, *r_ast, *p, *Ap, *s, *As, *tmp, *tmp2, *dinv;
    };

    // BiCGStab on vectors of type T (left preconditioned, restart on rho
    // breakdown, stop when ||r_k|| / ||M^{-1} b|| < tol).
    template <typename T>
    int bicgstab(const T *b, T *x, const Vecs &v, double tol, int maxit)
    {
        T *r = (T *)v.r, *r_ast = (T *)v.r_ast, *p = (T *)v.p, *Ap = (T *)v.Ap, *s = (T *)v.s, *As = (T *)v.As,
          *tmp = (T *)v.tmp, *tmp2 = (T *)v.tmp2;
        const T *dinv = (const T *)v.dinv;
        set_scalars(tol, maxit);
        prec<T>(b, tmp2, dinv);
        dots<T, 1>(DotArgs<T, 1>{{tmp2}, {tmp2}}, FIN_BNORM, false);
        auto init = [&]() {
            presidue<T>(x, b, r, tmp, dinv);
            hipLaunchKernelGGL((k_copy2<T>), g, KB, 0, c.stream, n, r, p, r_ast);
            MHFEM_HIP(hipGetLastError());
            dots<T, 1>(DotArgs<T, 1>{{r}, {r_ast}}, FIN_INIT, false);
        };
        init();
        const int every = std::max(1, c.cfg.check_every);
        for (;;) {
            poll();
            if (k.h_sc[S_DONE] != 0.0) break;
            if (k.h_sc[S_RESTART] != 0.0) {
                init();
                continue;
            }
            for (int q = 0; q < every; q++) {
                pmatvec<T, 1>(p, Ap, tmp, dinv, r_ast, FIN_ALPHA);
                hipLaunchKernelGGL((k_bicg_s<T>), g, KB, 0, c.stream, n, r, Ap, s, k.sc.p);
                pmatvec<T, 2>(s, As, tmp, dinv, s, FIN_OMEGA);
                hipLaunchKernelGGL((k_bicg_xr<T>), g, KB, 0, c.stream, n, x, p, s, As, r, r_ast, red);
                finish<2>(FIN_UPDATE, true);
                hipLaunchKernelGGL((k_bicg_p<T>), g, KB, 0, c.stream, n, r, p, Ap, k.sc.p);
                MHFEM_HIP(hipGetLastError());
            }
        }
        return (int)k.h_sc[S_IT];
    }

    // GMRES(m), FP64 only
    int gmres(const double *b, double *x, double tol, int maxit, int m, double *resid_out);
};

template <>
SellView<double> Solve::matrix<double>()
{
    return SellView<double>{c.rows, c.slice_ptr.p, c.slice_width.p, c.sell_col.p, c.sell_val.p};
}
template <>
SellView<float> Solve::matrix<float>()
{
    return SellView<float>{c.rows, c.slice_ptr.p, c.slice_width.p, c.sell_col.p, k.val32.p};
}

// w -= sum_k h[k] V_k  and  h2[k] = (V_k, w) for k <= j  (classical Gram-Schmidt pieces)
__global__ void __launch_bounds__(KB) k_gs_dots(int n, int j1, const double *__restrict__ V, const double *__restrict__ w,
                                                double *__restrict__ partials)
{
    // partial dots of w with V_0..V_{j1-1}, one row of partials per vector
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
    // one work-group per row of partials
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
        dots<double, 1>(DotArgs<double, 1>{{v}, {v}}, FIN_NONE, false);
        poll();
        return std::sqrt(k.h_sc[S_D0]);
    };
    set_scalars(tol, maxit);
    prec<double>(b, w, dinv);
    double b_norm = norm2(w);
    if (b_norm == 0.0) b_norm = 1.0;
    int it = 0;
    presidue<double>(x, b, V, tmp, dinv);
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
        presidue<double>(x, b, V, tmp, dinv);
        beta = norm2(V);
        resid = beta / b_norm;
    }
    MHFEM_HIP(hipStreamSynchronize(c.stream));
    *resid_out = resid;
    return it;
}

}  // namespace

// ---------------------------------------------------------------------------
// Preconditioner update (Solver.hpp:1237-1242): every step, from the freshly
// assembled matrix.
static void precond_update(Ctx &c, Krylov &k, Solve &S)
{
    const int n = k.n;
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL((k_diag_inv<double>), g, KB, 0, c.stream, n, c.sell_val.p, k.diag_pos.p, k.dinv.p);
    MHFEM_HIP(hipGetLastError());
    const bool mixed = c.cfg.precision == MHFEM_PRECISION_FP32_REFINED;
    if (mixed) {
        hipLaunchKernelGGL((k_convert<float, double>), grid_for(c.sell_size), KB, 0, c.stream, c.sell_size,
                           c.sell_val.p, k.val32.p);
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)n, k.dinv.p, k.dinv32.p);
        MHFEM_HIP(hipGetLastError());
    }
    if (c.cfg.preconditioner == MHFEM_PREC_ILU0) {
        Ilu0 &I = k.ilu;
        hipLaunchKernelGGL((k_gather<double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, c.sell_val.p,
                           c.csr_to_sell.p, I.lu.p);
        for (size_t l = 0; l + 1 < I.h_lptr.size(); l++) {
            const int cnt = I.h_lptr[l + 1] - I.h_lptr[l];
            hipLaunchKernelGGL(k_ilu_factor, (cnt + KB - 1) / KB, KB, 0, c.stream, I.lrows.p + I.h_lptr[l], cnt,
                               c.rowptr.p, c.csr_col.p, I.diag.p, I.lu.p);
        }
        if (mixed)
            hipLaunchKernelGGL((k_convert<float, double>), grid_for(c.nnz), KB, 0, c.stream, c.nnz, I.lu.p, I.lu32.p);
        MHFEM_HIP(hipGetLastError());
    }
    (void)S;
}

// ---------------------------------------------------------------------------
// FP64 helpers for callers outside the Krylov loop (learned initial guess):
// no convergence flag is consulted.
__global__ void __launch_bounds__(KB) k_spmv_plain(SellView<double> A, const double *__restrict__ x, double *__restrict__ y)
{
    for (int row = blockIdx.x * KB + threadIdx.x; row < A.rows; row += gridDim.x * KB) y[row] = sell_row(A, x, row);
}

void spmv64(Ctx &c, const double *x, double *y)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    hipLaunchKernelGGL(k_spmv_plain, grid_for(k.n), KB, 0, c.stream, S.matrix<double>(), x, y);
    MHFEM_HIP(hipGetLastError());
}

void residual64(Ctx &c, const double *x, const double *b, double *r)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    hipLaunchKernelGGL((k_residual<double>), grid_for(k.n), KB, 0, c.stream, S.matrix<double>(), x, b, r);
    MHFEM_HIP(hipGetLastError());
}

double dot64(Ctx &c, const double *x, const double *y)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    S.dots<double, 1>(DotArgs<double, 1>{{x}, {y}}, FIN_NONE, false);
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
    const double zero[NSLOTS] = {};
    MHFEM_HIP(hipMemcpyAsync(k.sc.p, zero, sizeof(zero), hipMemcpyHostToDevice, c.stream));
    if (precision == MHFEM_PRECISION_FP32_REFINED) {
        if (!k.val32.p) throw Error{MHFEM_E_STATE, "FP32 matrix not configured"};
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)k.n, c.arr[MHFEM_Z_iF].p, k.x32.p);
    }
    for (int rep = -2; rep < reps; rep++) {
        if (rep == 0) MHFEM_HIP(hipEventRecord(e0, c.stream));
        if (precision == MHFEM_PRECISION_FP32_REFINED)
            hipLaunchKernelGGL((k_spmv<float, false, 0>), g, KB, 0, c.stream, S.matrix<float>(), k.x32.p,
                               (const float *)nullptr, k.tmp32.p, (const float *)nullptr, S.red, 0);
        else
            hipLaunchKernelGGL((k_spmv<double, false, 0>), g, KB, 0, c.stream, S.matrix<double>(),
                               c.arr[MHFEM_Z_iF].p, (const double *)nullptr, k.tmp.p, (const double *)nullptr,
                               S.red, 0);
    }
    MHFEM_HIP(hipGetLastError());
    MHFEM_HIP(hipEventRecord(e1, c.stream));
    MHFEM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MHFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = reps > 0 ? ms / reps : 0.0;
}

void krylov_solve(Ctx &c, const double *b, double *x, mhfem_solve_stats *st)
{
    Krylov &k = *c.krylov;
    Solve S(c, k);
    const int n = k.n;
    const unsigned g = grid_for(n);
    hipEvent_t e0 = c.ev[3], e1 = c.ev[4], e2 = c.ev[5];
    if (c.timing) MHFEM_HIP(hipEventRecord(e0, c.stream));
    precond_update(c, k, S);
    if (c.timing) MHFEM_HIP(hipEventRecord(e1, c.stream));
    const double tol = c.cfg.convergence_residue;
    const int maxit = c.cfg.max_iterations;
    mhfem_solve_stats out{};
    if (c.guess && c.guess_on_solve) guess_before_solve(c, *c.guess, b, x);
    if (c.cfg.precision == MHFEM_PRECISION_FP64) {
        if (c.cfg.solver == MHFEM_SOLVER_GMRES) {
            double res = 0;
            out.iterations = S.gmres(b, x, tol, maxit, c.cfg.gmres_restart, &res);
            out.residue = res;
            out.converged = res < tol;
        } else {
            Solve::Vecs v{k.r.p, k.r_ast.p, k.p.p, k.Ap.p, k.s.p, k.As.p, k.tmp.p, k.tmp2.p, k.dinv.p};
            out.iterations = S.bicgstab<double>(b, x, v, tol, maxit);
            out.residue = k.h_sc[S_RESID];
            out.converged = out.residue < tol;
        }
        out.outer_iterations = 1;
    } else {
        // FP32 inner BiCGStab inside FP64 iterative refinement.  Outer test:
        // ||M^{-1}(b - A x)|| / ||M^{-1} b|| < tol, with M the FP32 preconditioner
        // applied to the FP64 residual rounded to FP32.
        float *rhs32 = k.rhs32.p, *z = k.b32.p;
        Solve::Vecs v{k.r32.p, k.r_ast32.p, k.p32.p, k.Ap32.p, k.s32.p, k.As32.p, k.tmp32.p, k.b32.p, k.dinv32.p};
        hipLaunchKernelGGL((k_convert<float, double>), g, KB, 0, c.stream, (long)n, b, rhs32);
        S.prec<float>(rhs32, z, k.dinv32.p);
        S.dots<float, 1>(DotArgs<float, 1>{{z}, {z}}, FIN_NONE, false);
        S.poll();
        double b_norm = std::sqrt(k.h_sc[S_D0]);
        if (b_norm == 0.0) b_norm = 1.0;
        const double inner_tol = c.cfg.inner_residue > 0 ? c.cfg.inner_residue : 1e-4;
        int total = 0, outer = 0;
        double res = 0;
        for (;;) {
            hipLaunchKernelGGL(k_residual_mixed, g, KB, 0, c.stream, S.matrix<double>(), x, b, rhs32);
            S.prec<float>(rhs32, z, k.dinv32.p);
            S.dots<float, 1>(DotArgs<float, 1>{{z}, {z}}, FIN_NONE, false);
            S.poll();
            res = std::sqrt(k.h_sc[S_D0]) / b_norm;
            if (!(res >= tol) || total >= maxit || outer >= 100) break;
            hipLaunchKernelGGL((k_fill_t<float>), g, KB, 0, c.stream, (long)n, k.x32.p, 0.0f);
            MHFEM_HIP(hipGetLastError());
            total += S.bicgstab<float>(rhs32, k.x32.p, v, inner_tol, std::max(1, maxit - total));
            hipLaunchKernelGGL(k_add_correction, g, KB, 0, c.stream, n, x, k.x32.p);
            MHFEM_HIP(hipGetLastError());
            outer++;
        }
        out.iterations = total;
        out.outer_iterations = outer;
        out.residue = res;
        out.converged = res < tol;
    }
    if (c.guess && c.guess_on_solve) guess_after_solve(c, *c.guess, x);
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

}  // namespace mhfem
