"""CPU restatement of the AMG preconditioner (the reference's Ginkgo "AMGX"
configuration, Solver.hpp:1025-1075): parallel graph matching aggregation,
Galerkin coarse matrices, block Jacobi, the V-cycle and the reuse rule
(Solver.hpp:1017, 1158-1160).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
library's hierarchy and V-cycle.  Ginkgo is not in the reference tree; the
steps restate its published PGM, block-Jacobi and multigrid algorithms, with
the same order of operations as the library where the order decides ties
(aggregation) or rounding (coarse values), so hierarchies agree exactly.
Pure-Python loops: small systems only.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

OMEGA = 0.9
MAX_LEVELS, MIN_COARSE_ROWS, PGM_ROUNDS, MAX_UNASSIGNED = 25, 4, 15, 0.1
PRE_SWEEPS, COARSEST_SWEEPS = 2, 4


def weights(A):
    """W = (|A| + |A^T|)/2 without the diagonal, and |a_ii| (as W's diagonal)."""
    A = sp.csr_matrix(A)
    W = (abs(A) + abs(A.T)).tocsr()
    W.sort_indices()
    W = W.multiply(0.5).tocsr()
    d = W.diagonal().copy()
    W.setdiag(0.0)
    W.eliminate_zeros()
    W.sort_indices()
    return W, d


def _strongest(W, d, agg, i, unagg):
    best, bw = -1, 0.0
    for q in range(W.indptr[i], W.indptr[i + 1]):
        j = W.indices[q]
        if (agg[j] == -1) != unagg:
            continue
        dm = max(d[i], d[j])
        w = W.data[q] / dm if dm > 0 else W.data[q]
        if w > bw or (w == bw and w > 0 and j > best):
            best, bw = j, w
    return best


def pgm(A):
    """Aggregate id of every row (0..nc-1, numbered by smallest member), nc."""
    W, d = weights(A)
    n = W.shape[0]
    agg = [-1] * n
    unassigned = n
    for _ in range(PGM_ROUNDS):
        if unassigned == 0:
            break
        snap = list(agg)
        strong = [-1] * n
        for i in range(n):
            if snap[i] != -1:
                continue
            u = _strongest(W, d, snap, i, True)
            if u >= 0:
                strong[i] = u
                continue
            a = _strongest(W, d, snap, i, False)
            if a >= 0:
                agg[i] = snap[a]
            else:
                strong[i] = i
        for i in range(n):
            if agg[i] != -1 or strong[i] < 0:
                continue
            j = strong[i]
            if strong[j] == i and i <= j and agg[j] == -1:
                agg[i] = i
                agg[j] = i
        unassigned = agg.count(-1)
        if unassigned <= MAX_UNASSIGNED * n:
            break
    if unassigned > 0:
        snap = list(agg)
        for i in range(n):
            if snap[i] != -1:
                continue
            a = _strongest(W, d, snap, i, False)
            agg[i] = snap[a] if a >= 0 else i
    ids = {}
    out = np.empty(n, np.int32)
    for i in range(n):
        if agg[i] not in ids:
            ids[agg[i]] = len(ids)
        out[i] = ids[agg[i]]
    return out, len(ids)


def galerkin(A, agg, nc):
    """P^T A P, summing the members' rows in ascending order, entries in row order."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    members = [[] for _ in range(nc)]
    for i, a in enumerate(agg):
        members[a].append(i)
    rows, cols, vals = [], [], []
    for I in range(nc):
        acc = {}
        for i in members[I]:
            for q in range(A.indptr[i], A.indptr[i + 1]):
                J = int(agg[A.indices[q]])
                acc[J] = acc.get(J, 0.0) + A.data[q]
        for J in sorted(acc):
            rows.append(I)
            cols.append(J)
            vals.append(acc[J])
    return sp.csr_matrix((vals, (rows, cols)), shape=(nc, nc))


def jacobi_blocks(A, bmax):
    """Block pointers (natural blocks of identical patterns, at most bmax rows,
    merged greedily up to bmax rows) and the inverse of every diagonal block."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    pat = [tuple(A.indices[A.indptr[r]:A.indptr[r + 1]]) for r in range(n)]
    nat = [0]
    for r in range(1, n + 1):
        if not (r < n and r - nat[-1] < bmax and pat[r] == pat[r - 1]):
            nat.append(r)
    bptr, cur = [0], 0
    for b in range(len(nat) - 1):
        size = nat[b + 1] - nat[b]
        if cur > 0 and cur + size <= bmax:
            cur += size
            bptr[-1] = nat[b + 1]
        else:
            cur = size
            bptr.append(nat[b + 1])
    dense = A.toarray() if n <= 4000 else None
    inv = []
    for b in range(len(bptr) - 1):
        r0, r1 = bptr[b], bptr[b + 1]
        blk = dense[r0:r1, r0:r1] if dense is not None else A[r0:r1, r0:r1].toarray()
        inv.append(np.linalg.inv(blk))
    return bptr, inv


class Hierarchy:
    def __init__(self, A, n_eq):
        A = sp.csr_matrix(A)
        A.sort_indices()
        self.A, self.agg, self.nc, self.blocks = [A], [], [], []
        while True:
            self.blocks.append(jacobi_blocks(self.A[-1], n_eq))
            if len(self.A) >= MAX_LEVELS or self.A[-1].shape[0] <= MIN_COARSE_ROWS:
                break
            agg, nc = pgm(self.A[-1])
            if nc >= self.A[-1].shape[0]:
                break
            self.agg.append(agg)
            self.nc.append(nc)
            self.A.append(galerkin(self.A[-1], agg, nc))

    def _binv(self, l, v):
        bptr, inv = self.blocks[l]
        out = np.empty_like(v)
        for b, B in enumerate(inv):
            out[bptr[b]:bptr[b + 1]] = B @ v[bptr[b]:bptr[b + 1]]
        return out

    def _sweep(self, l, b, x):
        return x + OMEGA * self._binv(l, b - self.A[l] @ x)

    def vcycle(self, b, l=0):
        """One V-cycle from a zero guess on level l."""
        x = OMEGA * self._binv(l, b)
        if l == len(self.A) - 1:
            for _ in range(COARSEST_SWEEPS - 1):
                x = self._sweep(l, b, x)
            return x
        for _ in range(PRE_SWEEPS - 1):
            x = self._sweep(l, b, x)
        r = b - self.A[l] @ x
        bc = np.bincount(self.agg[l], weights=r, minlength=self.nc[l])
        x = x + self.vcycle(bc, l + 1)[self.agg[l]]
        for _ in range(PRE_SWEEPS):
            x = self._sweep(l, b, x)
        return x


class ReuseRule:
    """Solver.hpp:1017, 1158-1160: rebuild unless last <= min(2 u, u + 5)."""

    def __init__(self):
        self.updated, self.last = 0, 1

    def rebuild(self):
        return not (self.last <= min(2 * self.updated, self.updated + 5))

    def after(self, iterations, rebuilt):
        self.last = iterations
        if rebuilt or self.updated == 0:
            self.updated = iterations
