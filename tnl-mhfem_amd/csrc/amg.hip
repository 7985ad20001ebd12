// amg.hip -- algebraic multigrid preconditioner: the configuration of the
// reference's Ginkgo branch (Solver.hpp:1025-1075, preconditioner "AMGX"):
//   * levels by parallel graph matching (PGM): pairs of unknowns joined along
//     their strongest mutual connection, deterministic, at most 15 matching
//     rounds or until at most 10 % of the unknowns are left, which then join
//     the aggregate of their strongest aggregated neighbour (alone if none);
//     piecewise-constant prolongation P, coarse matrix P^T A P; at most 25
//     levels, down to 4 rows;
//   * smoother: Richardson with relaxation 0.9 preconditioned by block
//     Jacobi (blocks of at most n rows), 2 sweeps before and after the
//     coarse correction; coarsest level: the same, 4 sweeps;
//   * one V-cycle from a zero initial guess per application;
//   * the hierarchy is reused while the last solve took at most
//     min(2 u, u + 5) iterations, u those of the solve right after the last
//     rebuild (Solver.hpp:1017, 1158-1160); the finest level always smooths
//     with the current matrix.
// Ginkgo itself is not in the reference tree: the steps are restated from its
// published PGM / block-Jacobi / multigrid algorithms.  The hierarchy is
// built on the host (setup is rare); the V-cycle runs on the device: per
// sweep a residual (the solver's block-SELL product on the finest level, CSR
// vector on the others) and a block update, then restriction and
// prolongation kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>

#include "common.hpp"

namespace mhfem {

namespace {

constexpr int AB = 256;  // work-group size of the V-cycle kernels
constexpr double OMEGA = 0.9;
constexpr int MAX_LEVELS = 25, MIN_COARSE_ROWS = 4, PGM_ROUNDS = 15;
constexpr double MAX_UNASSIGNED = 0.1;
constexpr int PRE_SWEEPS = 2, COARSEST_SWEEPS = 4;  // post-smoothing repeats the pre-smoother

struct HostCsr {
    int n = 0;
    std::vector<int> rp, col;
    std::vector<double> val;
};

// ---------------------------------------------------------------------------
// host setup

// W = (|A| + |A^T|) / 2 without its diagonal (the PGM weights), and |a_ii|
void weights(const HostCsr &A, HostCsr &W, std::vector<double> &diag)
{
    const int n = A.n;
    diag.assign(n, 0.0);
    // transpose
    std::vector<int> tp(n + 1, 0);
    for (int q = 0; q < A.rp[n]; q++) tp[A.col[q] + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    std::vector<int> tc(A.rp[n]);
    std::vector<double> tv(A.rp[n]);
    {
        std::vector<int> pos(tp.begin(), tp.end() - 1);
        for (int i = 0; i < n; i++)
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int p = pos[A.col[q]]++;
                tc[p] = i;
                tv[p] = A.val[q];
            }
    }
    W.n = n;
    W.rp.assign(n + 1, 0);
    W.col.clear();
    W.val.clear();
    W.col.reserve(A.rp[n]);
    W.val.reserve(A.rp[n]);
    for (int i = 0; i < n; i++) {
        int a = A.rp[i], b = tp[i];
        const int ae = A.rp[i + 1], be = tp[i + 1];
        while (a < ae || b < be) {
            const int ca = a < ae ? A.col[a] : INT32_MAX, cb = b < be ? tc[b] : INT32_MAX;
            const int cm = std::min(ca, cb);
            double v = 0.0;
            if (ca == cm) v += std::fabs(A.val[a++]);
            if (cb == cm) v += std::fabs(tv[b++]);
            if (cm == i) {
                diag[i] = 0.5 * v;
                continue;
            }
            W.col.push_back(cm);
            W.val.push_back(0.5 * v);
        }
        W.rp[i + 1] = (int)W.col.size();
    }
}

// strongest neighbour of row i under the rule: largest positive
// w_ij / max(d_i, d_j), ties to the larger index; among unaggregated
// neighbours (unagg = true) or aggregated ones; -1 if none
inline int strongest(const HostCsr &W, const std::vector<double> &d, const std::vector<int> &agg, int i,
                     bool unagg)
{
    int best = -1;
    double bw = 0.0;
    for (int q = W.rp[i]; q < W.rp[i + 1]; q++) {
        const int j = W.col[q];
        if ((agg[j] == -1) != unagg) continue;
        const double dm = std::max(d[i], d[j]);
        const double w = dm > 0.0 ? W.val[q] / dm : W.val[q];
        if (w > bw || (w == bw && w > 0.0 && j > best)) {
            best = j;
            bw = w;
        }
    }
    return best;
}

// PGM aggregates of A (ids 0 .. nc-1, numbered by their smallest member);
// returns nc
int pgm(const HostCsr &A, std::vector<int> &out)
{
    const int n = A.n;
    HostCsr W;
    std::vector<double> d;
    weights(A, W, d);
    std::vector<int> agg(n, -1), snap, strong(n);
    long unassigned = n;
    for (int round = 0; round < PGM_ROUNDS && unassigned > 0; round++) {
        snap = agg;
        for (int i = 0; i < n; i++) {
            strong[i] = -1;
            if (snap[i] != -1) continue;
            const int u = strongest(W, d, snap, i, true);
            if (u >= 0) {
                strong[i] = u;
                continue;
            }
            const int a = strongest(W, d, snap, i, false);
            if (a >= 0) agg[i] = snap[a];  // no free neighbour: join an existing aggregate
            else strong[i] = i;            // isolated: an aggregate of its own
        }
        for (int i = 0; i < n; i++) {
            if (agg[i] != -1 || strong[i] < 0) continue;
            const int j = strong[i];
            if (strong[j] == i && i <= j && agg[j] == -1) {
                agg[i] = i;
                agg[j] = i;
            }
        }
        unassigned = std::count(agg.begin(), agg.end(), -1);
        if (unassigned <= MAX_UNASSIGNED * n) break;
    }
    if (unassigned > 0) {
        snap = agg;
        for (int i = 0; i < n; i++) {
            if (snap[i] != -1) continue;
            const int a = strongest(W, d, snap, i, false);
            agg[i] = a >= 0 ? snap[a] : i;
        }
    }
    // renumber by representative
    std::vector<int> id(n, -1);
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (id[agg[i]] < 0) id[agg[i]] = nc++;
    out.resize(n);
    for (int i = 0; i < n; i++) out[i] = id[agg[i]];
    return nc;
}

// P^T A P: coarse row I sums the rows of its members (ascending), each
// entry added to the coarse column of its fine column in row order
void galerkin(const HostCsr &A, const std::vector<int> &agg, int nc, const std::vector<int> &pt_ptr,
              const std::vector<int> &pt_idx, HostCsr &Ac)
{
    Ac.n = nc;
    Ac.rp.assign(nc + 1, 0);
    Ac.col.clear();
    Ac.val.clear();
    std::vector<double> acc(nc, 0.0);
    std::vector<int> mark(nc, -1), cols;
    for (int I = 0; I < nc; I++) {
        cols.clear();
        for (int m = pt_ptr[I]; m < pt_ptr[I + 1]; m++) {
            const int i = pt_idx[m];
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int J = agg[A.col[q]];
                if (mark[J] != I) {
                    mark[J] = I;
                    acc[J] = 0.0;
                    cols.push_back(J);
                }
                acc[J] += A.val[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int J : cols) {
            Ac.col.push_back(J);
            Ac.val.push_back(acc[J]);
        }
        Ac.rp[I + 1] = (int)Ac.col.size();
    }
}

// block-Jacobi blocks: consecutive rows with the same column pattern form
// natural blocks (at most bmax rows), which are merged greedily while the
// merged block has at most bmax rows; the inverse of each diagonal block,
// row-major in 9 slots, by Gauss-Jordan elimination with partial pivoting
void jacobi_blocks(const HostCsr &A, int bmax, std::vector<int> &bptr, std::vector<double> &binv)
{
    const int n = A.n;
    std::vector<int> nat(1, 0);
    for (int r = 1; r <= n; r++) {
        bool same = r < n && r - nat.back() < bmax;
        if (same) {
            const int la = A.rp[r] - A.rp[r - 1], lb = A.rp[r + 1] - A.rp[r];
            same = la == lb && std::equal(A.col.begin() + A.rp[r - 1], A.col.begin() + A.rp[r], A.col.begin() + A.rp[r]);
        }
        if (!same) nat.push_back(r);
    }
    bptr.assign(1, 0);
    int cur = 0;
    for (size_t b = 0; b + 1 < nat.size(); b++) {
        const int size = nat[b + 1] - nat[b];
        if (cur > 0 && cur + size <= bmax) {
            cur += size;
            bptr.back() = nat[b + 1];
        } else {
            cur = size;
            bptr.push_back(nat[b + 1]);
        }
    }
    const int nb = (int)bptr.size() - 1;
    binv.assign((size_t)nb * 9, 0.0);
    for (int b = 0; b < nb; b++) {
        const int r0 = bptr[b], bs = bptr[b + 1] - r0;
        double M[3][6] = {};
        for (int r = 0; r < bs; r++) {
            for (int q = A.rp[r0 + r]; q < A.rp[r0 + r + 1]; q++) {
                const int cc = A.col[q] - r0;
                if (cc >= 0 && cc < bs) M[r][cc] = A.val[q];
            }
            M[r][bs + r] = 1.0;
        }
        for (int k = 0; k < bs; k++) {
            int p = k;
            for (int r = k + 1; r < bs; r++)
                if (std::fabs(M[r][k]) > std::fabs(M[p][k])) p = r;
            if (M[p][k] == 0.0) throw Error{MHFEM_E_BREAKDOWN, "AMG: singular block-Jacobi block"};
            if (p != k)
                for (int c = 0; c < 2 * bs; c++) std::swap(M[k][c], M[p][c]);
            const double piv = M[k][k];
            for (int c = 0; c < 2 * bs; c++) M[k][c] /= piv;
            for (int r = 0; r < bs; r++) {
                if (r == k) continue;
                const double f = M[r][k];
                if (f == 0.0) continue;
                for (int c = 0; c < 2 * bs; c++) M[r][c] -= f * M[k][c];
            }
        }
        for (int r = 0; r < bs; r++)
            for (int c = 0; c < bs; c++) binv[(size_t)b * 9 + r * 3 + c] = M[r][bs + c];
    }
}

// ---------------------------------------------------------------------------
// device V-cycle

// sc: flags of the single segment (skip once it has halted) or nullptr
__device__ __forceinline__ bool amg_halted(const double *sc) { return sc && halted(sc); }

// x = omega B^{-1} b (the first sweep from a zero guess)
__global__ void __launch_bounds__(AB) k_amg_bj_init(int nb, const int *__restrict__ bptr,
                                                    const double *__restrict__ binv, const double *__restrict__ b,
                                                    double *__restrict__ x, const double *sc)
{
    const int k = blockIdx.x * AB + threadIdx.x;
    if (k >= nb || amg_halted(sc)) return;
    const int r0 = bptr[k], bs = bptr[k + 1] - r0;
    const double *B = binv + (size_t)k * 9;
    double v[3];
    for (int r = 0; r < bs; r++) v[r] = b[r0 + r];
    for (int r = 0; r < bs; r++) {
        double s = 0.0;
        for (int c = 0; c < bs; c++) s += B[r * 3 + c] * v[c];
        x[r0 + r] = OMEGA * s;
    }
}

// r = b - A x on a coarse level: L lanes per row (CSR vector), partial sums
// combined by a fixed butterfly
template <int L>
__global__ void __launch_bounds__(AB) k_amg_residual(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                                     const double *__restrict__ val, const double *__restrict__ x,
                                                     const double *__restrict__ b, double *__restrict__ r,
                                                     const double *sc)
{
    if (amg_halted(sc)) return;
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    const int row = (int)(t / L), lane = (int)(t % L);
    double s = 0.0;
    if (row < n)
        for (int q = rp[row] + lane; q < rp[row + 1]; q += L) s += val[q] * x[col[q]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (row < n && lane == 0) r[row] = b[row] - s;
}

// x += omega B^{-1} r, block by block
__global__ void __launch_bounds__(AB) k_amg_bj_update(int nb, const int *__restrict__ bptr,
                                                      const double *__restrict__ binv, const double *__restrict__ r,
                                                      double *__restrict__ x, const double *sc)
{
    const int k = blockIdx.x * AB + threadIdx.x;
    if (k >= nb || amg_halted(sc)) return;
    const int r0 = bptr[k], bs = bptr[k + 1] - r0;
    const double *B = binv + (size_t)k * 9;
    double v[3];
    for (int i = 0; i < bs; i++) v[i] = r[r0 + i];
    for (int i = 0; i < bs; i++) {
        double s = 0.0;
        for (int c = 0; c < bs; c++) s += B[i * 3 + c] * v[c];
        x[r0 + i] += OMEGA * s;
    }
}

// bc[I] = sum of r over the members of aggregate I
__global__ void __launch_bounds__(AB) k_amg_restrict(int nc, const int *__restrict__ pt_ptr,
                                                     const int *__restrict__ pt_idx, const double *__restrict__ r,
                                                     double *__restrict__ bc, const double *sc)
{
    const int I = blockIdx.x * AB + threadIdx.x;
    if (I >= nc || amg_halted(sc)) return;
    double acc = 0.0;
    for (int m = pt_ptr[I]; m < pt_ptr[I + 1]; m++) acc += r[pt_idx[m]];
    bc[I] = acc;
}

// x += P xc
__global__ void __launch_bounds__(AB) k_amg_prolong(int n, const int *__restrict__ agg,
                                                    const double *__restrict__ xc, double *__restrict__ x,
                                                    const double *sc)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n || amg_halted(sc)) return;
    x[i] += xc[agg[i]];
}

inline unsigned agrid(long n) { return (unsigned)((n + AB - 1) / AB); }

// every index the V-cycle kernels follow stays inside its array
void check_level(const HostCsr &Ac, const std::vector<int> &agg, int nc, int nf)
{
    bool ok = (int)Ac.rp.size() == nc + 1 && Ac.rp[0] == 0 && (int)agg.size() == nf;
    for (int I = 0; ok && I < nc; I++) ok = Ac.rp[I] <= Ac.rp[I + 1];
    ok = ok && Ac.rp[nc] == (int)Ac.col.size();
    for (size_t q = 0; ok && q < Ac.col.size(); q++) ok = Ac.col[q] >= 0 && Ac.col[q] < nc;
    for (int i = 0; ok && i < nf; i++) ok = agg[i] >= 0 && agg[i] < nc;
    if (!ok) throw Error{MHFEM_E_STATE, "AMG: inconsistent coarse level"};
}

}  // namespace

struct AmgLevel {
    int n = 0, nc = 0, nb = 0;
    long nnz = 0;
    DBuf<int> rp, col, agg, pt_ptr, pt_idx, bptr;  // rp, col, val: coarse levels (the finest uses the solver's matrix)
    DBuf<double> val, binv, b, x, r;
    std::vector<int> h_agg;  // for inspection
};

struct Amg {
    std::vector<std::unique_ptr<AmgLevel>> L;  // levels stay in place while more are added
    long long updated_iters = 0, last_iters = 1;
    bool rebuilt = false;
    int rebuilds = 0;
};

Amg *amg_create() { return new Amg(); }
void amg_destroy(Amg *a) { delete a; }

static void build(Ctx &c, Amg &a)
{
    a.L.clear();
    a.L.push_back(std::make_unique<AmgLevel>());
    {
        AmgLevel &f = *a.L[0];
        f.n = c.rows;
        f.nnz = c.nnz;
    }
    // the finest level's values in CSR order (unit diagonal included), for the host setup
    HostCsr A;
    A.n = c.rows;
    A.rp = c.h_rowptr;
    A.col = c.h_col;
    A.val.resize(c.nnz);
    {
        DBuf<double> tmp;
        tmp.alloc(std::max<long>(c.nnz, 1));
        launch_csr_gather(c, tmp.p);
        MHFEM_HIP(hipMemcpyAsync(A.val.data(), tmp.p, c.nnz * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    }
    for (;;) {
        AmgLevel &lv = *a.L.back();
        std::vector<int> bptr;
        std::vector<double> binv;
        jacobi_blocks(A, c.n, bptr, binv);
        lv.nb = (int)bptr.size() - 1;
        if (bptr.front() != 0 || bptr.back() != A.n) throw Error{MHFEM_E_STATE, "AMG: inconsistent Jacobi blocks"};
        for (int b = 0; b < lv.nb; b++)
            if (bptr[b + 1] - bptr[b] < 1 || bptr[b + 1] - bptr[b] > 3)
                throw Error{MHFEM_E_STATE, "AMG: inconsistent Jacobi blocks"};
        lv.bptr.alloc(bptr.size());
        lv.binv.alloc(binv.size());
        MHFEM_HIP(hipMemcpy(lv.bptr.p, bptr.data(), bptr.size() * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(lv.binv.p, binv.data(), binv.size() * sizeof(double), hipMemcpyHostToDevice));
        lv.r.alloc(A.n);
        if ((int)a.L.size() >= MAX_LEVELS || A.n <= MIN_COARSE_ROWS) break;
        std::vector<int> agg;
        const int nc = pgm(A, agg);
        if (nc >= A.n) break;  // no coarsening left
        std::vector<int> pt_ptr(nc + 1, 0), pt_idx(A.n);
        for (int i = 0; i < A.n; i++) pt_ptr[agg[i] + 1]++;
        for (int I = 0; I < nc; I++) pt_ptr[I + 1] += pt_ptr[I];
        {
            std::vector<int> pos(pt_ptr.begin(), pt_ptr.end() - 1);
            for (int i = 0; i < A.n; i++) pt_idx[pos[agg[i]]++] = i;
        }
        HostCsr Ac;
        galerkin(A, agg, nc, pt_ptr, pt_idx, Ac);
        check_level(Ac, agg, nc, A.n);
        lv.nc = nc;
        lv.h_agg = agg;
        lv.agg.alloc(A.n);
        lv.pt_ptr.alloc(nc + 1);
        lv.pt_idx.alloc(A.n);
        MHFEM_HIP(hipMemcpy(lv.agg.p, agg.data(), A.n * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(lv.pt_ptr.p, pt_ptr.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(lv.pt_idx.p, pt_idx.data(), A.n * sizeof(int), hipMemcpyHostToDevice));
        a.L.push_back(std::make_unique<AmgLevel>());
        AmgLevel &cl = *a.L.back();
        cl.n = nc;
        cl.nnz = (long)Ac.col.size();
        cl.rp.alloc(nc + 1);
        cl.col.alloc(std::max<long>(cl.nnz, 1));
        cl.val.alloc(std::max<long>(cl.nnz, 1));
        MHFEM_HIP(hipMemcpy(cl.rp.p, Ac.rp.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(cl.col.p, Ac.col.data(), cl.nnz * sizeof(int), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(cl.val.p, Ac.val.data(), cl.nnz * sizeof(double), hipMemcpyHostToDevice));
        cl.b.alloc(nc);
        cl.x.alloc(nc);
        A = std::move(Ac);
    }
    a.rebuilds++;
}

// Preconditioner update before a solve (Solver.hpp:1012-1017): rebuild unless
// the reuse rule holds.  The finest level's residuals always use the solver's
// current matrix.
void amg_update(Ctx &c, Amg &a)
{
    const bool reuse = !a.L.empty() && a.last_iters <= std::min(2 * a.updated_iters, a.updated_iters + 5);
    a.rebuilt = !reuse;
    if (!reuse) build(c, a);
}

// after the solve (Solver.hpp:1158-1160)
void amg_after_solve(Amg &a, long long iterations)
{
    a.last_iters = iterations;
    if (a.rebuilt || a.updated_iters == 0) a.updated_iters = iterations;
}

// r = b - A x on level l: the solver's block-SELL matrix on the finest level,
// CSR vector on the coarse ones (lanes per row from the mean row length)
static void level_residual(Ctx &c, AmgLevel &v, int l, const double *x, const double *b, double *r,
                           const double *sc)
{
    if (l == 0) {
        residual64_halt(c, x, b, r, sc);
        return;
    }
    const long avg = v.nnz / std::max(1, v.n);
    auto go = [&](auto LC) {
        constexpr int L = decltype(LC)::value;
        hipLaunchKernelGGL((k_amg_residual<L>), agrid((long)v.n * L), AB, 0, c.stream, v.n, v.rp.p, v.col.p, v.val.p,
                           x, b, r, sc);
    };
    if (avg <= 6) go(std::integral_constant<int, 4>{});
    else if (avg <= 12) go(std::integral_constant<int, 8>{});
    else if (avg <= 24) go(std::integral_constant<int, 16>{});
    else go(std::integral_constant<int, 32>{});
    MHFEM_HIP(hipGetLastError());
}

// one V-cycle on level l from a zero guess: x = M_l^{-1} b (x and b distinct)
static void vcycle(Ctx &c, Amg &a, int l, const double *b, double *x, const double *sc)
{
    AmgLevel &v = *a.L[l];
    const bool last = l + 1 == (int)a.L.size();
    auto sweep = [&]() {
        level_residual(c, v, l, x, b, v.r.p, sc);
        hipLaunchKernelGGL(k_amg_bj_update, agrid(v.nb), AB, 0, c.stream, v.nb, v.bptr.p, v.binv.p, v.r.p, x, sc);
    };
    // the first Richardson step from zero: x = omega B^{-1} b
    hipLaunchKernelGGL(k_amg_bj_init, agrid(v.nb), AB, 0, c.stream, v.nb, v.bptr.p, v.binv.p, b, x, sc);
    if (last) {
        for (int s = 1; s < COARSEST_SWEEPS; s++) sweep();
        MHFEM_HIP(hipGetLastError());
        return;
    }
    for (int s = 1; s < PRE_SWEEPS; s++) sweep();
    AmgLevel &w = *a.L[l + 1];
    level_residual(c, v, l, x, b, v.r.p, sc);
    hipLaunchKernelGGL(k_amg_restrict, agrid(v.nc), AB, 0, c.stream, v.nc, v.pt_ptr.p, v.pt_idx.p, v.r.p, w.b.p, sc);
    vcycle(c, a, l + 1, w.b.p, w.x.p, sc);
    hipLaunchKernelGGL(k_amg_prolong, agrid(v.n), AB, 0, c.stream, v.n, v.agg.p, w.x.p, x, sc);
    for (int s = 0; s < PRE_SWEEPS; s++) sweep();
    MHFEM_HIP(hipGetLastError());
}

// y = M^{-1} x: one V-cycle from a zero guess
void amg_apply(Ctx &c, Amg &a, const double *x, double *y, const double *sc)
{
    if (a.L.empty()) throw Error{MHFEM_E_STATE, "AMG hierarchy not built"};
    if (x == y) throw Error{MHFEM_E_STATE, "AMG: the V-cycle needs distinct input and output vectors"};
    vcycle(c, a, 0, x, y, sc);
}

int amg_num_levels(const Amg &a) { return (int)a.L.size(); }
int amg_rebuilds(const Amg &a) { return a.rebuilds; }
void amg_level_info(const Amg &a, int l, int *rows, long *nnz, int32_t *agg)
{
    const AmgLevel &v = *a.L.at(l);
    if (rows) *rows = v.n;
    if (nnz) *nnz = v.nnz;
    if (agg && !v.h_agg.empty()) std::memcpy(agg, v.h_agg.data(), v.h_agg.size() * sizeof(int32_t));
}

}  // namespace mhfem
