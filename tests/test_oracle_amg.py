"""Pins of the AMG oracle (oracle/amg_oracle.py), the checker of the
library's AMG preconditioner: Galerkin products against scipy, the matching
rule on hand-derived chains, block-Jacobi inverses, the V-cycle against its
operator written as a matrix recursion, V-cycle convergence on SPD systems,
and the reuse rule (Solver.hpp:1017, 1158-1160) on a hand-worked sequence.
Ginkgo is not in the reference tree, so this restatement is pinned by these
properties only (parity unpinned against Ginkgo itself)."""
import numpy as np
import pytest
import scipy.sparse as sp

import amg_oracle as amg
import pyoracle
from mhfem_amd import problems


def chain(weights):
    """Tridiagonal M-matrix with the given off-diagonal couplings and a unit diagonal."""
    n = len(weights) + 1
    off = -np.asarray(weights, float)
    A = sp.diags([off, np.ones(n), off], [-1, 0, 1], format="csr")
    return A


def test_pgm_on_a_chain_by_hand():
    # couplings 1,3,1,3,1,3: rows (1,2), (3,4), (5,6) are mutual strongest;
    # row 0 then has no free neighbour and joins its strongest aggregated one
    agg, nc = amg.pgm(chain([1, 3, 1, 3, 1, 3]))
    assert nc == 3
    assert agg.tolist() == [0, 0, 0, 1, 1, 2, 2]
    # equal couplings: ties go to the larger index, so row i points to i+1 and
    # only the last pair (3, 4) is mutual in the first round; then (1, 2)
    # (row 1 prefers 2, row 2 has only 1 left); row 0 joins row 1's aggregate
    agg, nc = amg.pgm(chain([1, 1, 1, 1]))
    assert agg.tolist() == [0, 0, 0, 1, 1] and nc == 2


def test_pgm_isolated_rows_are_their_own_aggregates():
    A = sp.identity(5, format="csr")
    agg, nc = amg.pgm(A)
    assert nc == 5 and agg.tolist() == [0, 1, 2, 3, 4]


def _mhfem_matrix(kind="adr3d"):
    pr = problems.adr_3d(s=3, n=2, preconditioner="diagonal") if kind == "adr3d" else problems.heat_2d(s=8)
    o = pyoracle.Oracle(pr)
    o.pre_iterate(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    rp, col, val, rhs = o.csr()
    return sp.csr_matrix((val.copy(), col.copy(), rp.copy())), pr.n


@pytest.mark.parametrize("kind", ["adr3d", "heat2d"])
def test_hierarchy_properties(kind):
    A, n = _mhfem_matrix(kind)
    H = amg.Hierarchy(A, n)
    assert len(H.A) >= 3 and H.A[-1].shape[0] <= amg.MIN_COARSE_ROWS or len(H.A) == amg.MAX_LEVELS
    for l, agg in enumerate(H.agg):
        nf, nc = H.A[l].shape[0], H.nc[l]
        # ids 0..nc-1, numbered by their smallest member
        first = np.full(nc, -1)
        for i, a in enumerate(agg):
            if first[a] < 0:
                first[a] = i
        assert np.all(np.diff(first) > 0)
        assert nc < nf and nc >= nf // 4
        # Galerkin product equals P^T A P
        P = sp.csr_matrix((np.ones(nf), (np.arange(nf), agg)), shape=(nf, nc))
        ref = (P.T @ H.A[l] @ P).toarray()
        assert np.abs(H.A[l + 1].toarray() - ref).max() <= 1e-13 * np.abs(ref).max()
    for l, (bptr, inv) in enumerate(H.blocks):
        D = H.A[l].toarray()
        assert bptr[0] == 0 and bptr[-1] == D.shape[0]
        assert max(np.diff(bptr)) <= n
        for b, B in enumerate(inv):
            r0, r1 = bptr[b], bptr[b + 1]
            assert np.allclose(B @ D[r0:r1, r0:r1], np.eye(r1 - r0), atol=1e-12)


def test_vcycle_equals_its_matrix_recursion():
    A, n = _mhfem_matrix("adr3d")
    H = amg.Hierarchy(A, n)

    def smoother_matrix(l):
        bptr, inv = H.blocks[l]
        S = np.zeros((bptr[-1], bptr[-1]))
        for b, B in enumerate(inv):
            S[bptr[b]:bptr[b + 1], bptr[b]:bptr[b + 1]] = amg.OMEGA * B
        return S

    def M(l):
        Al = H.A[l].toarray()
        S = smoother_matrix(l)
        I = np.eye(Al.shape[0])
        X = S.copy()
        if l == len(H.A) - 1:
            for _ in range(amg.COARSEST_SWEEPS - 1):
                X = X + S @ (I - Al @ X)
            return X
        for _ in range(amg.PRE_SWEEPS - 1):
            X = X + S @ (I - Al @ X)
        P = np.zeros((Al.shape[0], H.nc[l]))
        P[np.arange(Al.shape[0]), H.agg[l]] = 1.0
        X = X + P @ M(l + 1) @ P.T @ (I - Al @ X)
        for _ in range(amg.PRE_SWEEPS):
            X = X + S @ (I - Al @ X)
        return X

    b = np.random.default_rng(1).standard_normal(A.shape[0])
    ref = M(0) @ b
    got = H.vcycle(b)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_vcycle_converges_on_spd_system():
    A, n = _mhfem_matrix("heat2d")
    A = (A + A.T) * 0.5  # the heat system is diagonally scaled SPD: symmetrise the scaling away
    H = amg.Hierarchy(A, n)
    Ad = A.toarray()
    E = np.eye(A.shape[0]) - np.column_stack([H.vcycle(Ad[:, j]) for j in range(A.shape[0])])
    rho = max(abs(np.linalg.eigvals(E)))
    assert rho < 0.9, rho


def test_reuse_rule_by_hand():
    r = amg.ReuseRule()
    pattern = []
    for its in [10, 12, 21, 9, 16, 15, 14]:
        rebuild = r.rebuild()
        pattern.append(rebuild)
        r.after(its, rebuild)
    # 1st: no hierarchy (last 1 > min(0, 5)); then u = 10: 12 <= 15 reuse; 21 > 15
    # rebuild after it; u = 9: 16 > 14 rebuild after it; u = 15: 14 <= 20
    assert pattern == [True, False, False, True, False, True, False]
