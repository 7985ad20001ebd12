"""CPU restatement of the ILU-ISAI preconditioners (the reference's Ginkgo
"ILU_ISAI" and "PARILU_ISAI" configurations, Solver.hpp:1089-1130): ILU(0)
factors L (unit lower) and U, exact or by ParILU sweeps, each replaced by an
incomplete sparse approximate inverse whose pattern is that of the factor's
second power, applied as y = M_U (M_L x).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
library's ILU-ISAI preconditioner.  Ginkgo is not in the reference tree; this
restates the published method (Anzt, Huckle, Braeckle, Dongarra 2018,
"Incomplete sparse approximate inverses for parallel preconditioning"):
row i of M solves (M T)_{i,j} = delta_ij for every j in the pattern J_i of
row i, i.e. the small triangular system T[J_i, J_i]^T m = e_i.  Every row is
solved exactly (Ginkgo solves rows longer than 32 approximately, through an
"excess" system).  Pure-Python loops: small systems only.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

SPARSITY_POWER = 2  # Solver.hpp:1111


def ilu0(A):
    """ILU(0) of a CSR matrix with sorted columns and a full diagonal (IKJ
    form, Saad Alg. 10.4): returns (L, U), L unit lower triangular with its
    diagonal stored, U upper triangular, both on A's pattern."""
    A = sp.csr_matrix(A, dtype=float, copy=True)
    A.sort_indices()
    rp, col, lu = A.indptr, A.indices, A.data
    n = A.shape[0]
    diag = np.array([rp[i] + np.searchsorted(col[rp[i]:rp[i + 1]], i) for i in range(n)])
    for i in range(n):
        pos = {col[q]: q for q in range(rp[i], rp[i + 1])}
        for q in range(rp[i], diag[i]):
            k = col[q]
            lu[q] /= lu[diag[k]]
            for t in range(diag[k] + 1, rp[k + 1]):
                p = pos.get(col[t])
                if p is not None:
                    lu[p] -= lu[q] * lu[t]
    LU = sp.csr_matrix((lu, col, rp), shape=A.shape)
    L = sp.tril(LU, -1, format="csr") + sp.identity(n, format="csr")
    U = sp.triu(LU, 0, format="csr")
    L.sort_indices()
    U.sort_indices()
    return L, U


PARILU_SWEEPS = 5  # synchronous sweeps of the ParILU fixed-point iteration


def parilu(A, sweeps=PARILU_SWEEPS):
    """ParILU (Chow & Patel 2015): the ILU(0) equations on A's pattern S,
        l_ij u_jj + sum_{k<j} l_ik u_kj = a_ij   (i > j),
        u_ij + sum_{k<i} l_ik u_kj = a_ij         (i <= j),
    solved by fixed-point sweeps from l_ij = a_ij / a_jj, u_ij = a_ij.  Each
    sweep evaluates every entry from the previous sweep's values (synchronous
    sweeps; Ginkgo updates in place, asynchronously).  Returns (L, U) as
    ilu0 does; the fixed point is ILU(0)."""
    A = sp.csr_matrix(A, dtype=float, copy=True)
    A.sort_indices()
    rp, col, a = A.indptr, A.indices, A.data
    n = A.shape[0]
    diag = np.array([rp[i] + np.searchsorted(col[rp[i]:rp[i + 1]], i) for i in range(n)])
    pos = [{col[q]: q for q in range(rp[i], rp[i + 1])} for i in range(n)]
    lu = a.copy()
    for i in range(n):
        for q in range(rp[i], diag[i]):
            lu[q] = a[q] / a[diag[col[q]]]
    for _ in range(sweeps):
        new = np.empty_like(lu)
        for i in range(n):
            for q in range(rp[i], rp[i + 1]):
                j = col[q]
                m = min(i, j)
                acc = a[q]
                for q2 in range(rp[i], rp[i + 1]):
                    k = col[q2]
                    if k >= m:
                        break
                    p = pos[k].get(j)
                    if p is not None:
                        acc -= lu[q2] * lu[p]
                new[q] = acc / lu[diag[j]] if i > j else acc
        lu = new
    LU = sp.csr_matrix((lu, col, rp), shape=A.shape)
    L = sp.tril(LU, -1, format="csr") + sp.identity(n, format="csr")
    U = sp.triu(LU, 0, format="csr")
    L.sort_indices()
    U.sort_indices()
    return L, U


def pattern_power(T, k):
    """Structural pattern of T^k (boolean CSR, sorted)."""
    P = sp.csr_matrix((np.ones(T.nnz), T.indices, T.indptr), shape=T.shape)
    R = P.copy()
    for _ in range(k - 1):
        R = (R @ P).tocsr()
    R.data[:] = 1.0
    R.sort_indices()
    return R


def isai(T, lower, k=SPARSITY_POWER):
    """Incomplete sparse approximate inverse of the triangular factor T on the
    pattern of T^k: row i solves T[J, J]^T m = e_i."""
    T = sp.csr_matrix(T)
    P = pattern_power(T, k)
    Td = T.toarray() if T.shape[0] <= 4000 else None
    vals = np.zeros(P.nnz)
    for i in range(T.shape[0]):
        J = P.indices[P.indptr[i]:P.indptr[i + 1]]
        sub = Td[np.ix_(J, J)] if Td is not None else T[J][:, J].toarray()
        e = np.zeros(J.size)
        e[np.searchsorted(J, i)] = 1.0
        vals[P.indptr[i]:P.indptr[i + 1]] = np.linalg.solve(sub.T, e)
    M = sp.csr_matrix((vals, P.indices.copy(), P.indptr.copy()), shape=T.shape)
    return M


def isai_rows(T, rows, lower, k=SPARSITY_POWER):
    """Rows `rows` of the ISAI of T, each as (pattern, values); for spot checks
    of large systems (T in CSR with sorted columns)."""
    T = sp.csr_matrix(T)
    out = []
    for i in rows:
        J = T.indices[T.indptr[i]:T.indptr[i + 1]]
        for _ in range(k - 1):
            J = np.unique(np.concatenate([T.indices[T.indptr[t]:T.indptr[t + 1]] for t in J]))
        sub = T[J][:, J].toarray()
        e = np.zeros(J.size)
        e[np.searchsorted(J, i)] = 1.0
        out.append((J, np.linalg.solve(sub.T, e)))
    return out


class IluIsai:
    """y = M_U (M_L x) with M_L, M_U the ISAIs of A's ILU(0) factors (or of
    its ParILU factors with `sweeps` synchronous sweeps)."""

    def __init__(self, A, k=SPARSITY_POWER, sweeps=None):
        self.L, self.U = ilu0(A) if sweeps is None else parilu(A, sweeps)
        self.ML = isai(self.L, True, k)
        self.MU = isai(self.U, False, k)

    def __call__(self, x):
        return self.MU @ (self.ML @ x)


def bicgstab(A, b, x0, M, tol, maxit=10000):
    """Left-preconditioned BiCGStab with the library's stopping test
    ||M(b - A x)|| / ||M b|| < tol and restart on rho breakdown; M a callable."""
    x = x0.astype(float).copy()
    bn = np.linalg.norm(M(b)) or 1.0
    r = M(b - A @ x)
    rs, p = r.copy(), r.copy()
    rho = r @ rs
    rs_sq = rho
    res, it = np.sqrt(rho) / bn, 0
    eps2 = np.finfo(float).eps ** 2
    while res >= tol and it < maxit and np.isfinite(res):
        Ap = M(A @ p)
        den = Ap @ rs
        alpha = rho / den if den != 0 else np.inf
        if not (abs(den) > 0) or not np.isfinite(alpha):
            r = M(b - A @ x)
            rs, p = r.copy(), r.copy()
            rho = rs_sq = r @ rs
            it += 1
            res = np.linalg.norm(r) / bn
            continue
        s = r - alpha * Ap
        As = M(A @ s)
        aa = As @ As
        omega = (As @ s) / aa if aa > 0 else 0.0
        x += alpha * p + omega * s
        r = s - omega * As
        rho_old, rho = rho, r @ rs
        if not (abs(rho) >= eps2 * rs_sq) or omega == 0.0:
            r = M(b - A @ x)
            rs, p = r.copy(), r.copy()
            rho = rs_sq = r @ rs
        else:
            beta = (rho / rho_old) * (alpha / omega)
            p = r + beta * (p - omega * Ap)
        it += 1
        res = np.linalg.norm(r) / bn
    return x, it, res
