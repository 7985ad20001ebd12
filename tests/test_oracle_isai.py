"""Pins of the ILU-ISAI oracle (oracle/isai_oracle.py), the checker of the
library's ILU-ISAI preconditioner.  Ginkgo is not in the reference tree, so
the restatement is pinned by what the mathematics fixes:
* ILU(0) agrees bit for bit with the C oracle's independent ILU(0), and L U
  reproduces A on A's pattern (the defining property of ILU(0));
* on a tridiagonal matrix ILU(0) is the exact LU, and the ISAI of its
  bidiagonal factors on the power-2 pattern equals the exact inverse there
  (closed form: -l_i and l_i l_{i-1} below the unit diagonal);
* each row of an ISAI satisfies (M T)_{i,j} = delta_ij on its pattern, with
  the product taken by scipy;
* when the pattern covers the whole triangle, the ISAI is the exact inverse;
* BiCGStab preconditioned with it needs about as many iterations as with the
  exact ILU(0) solves, and far fewer than without a preconditioner."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import isai_oracle as isai
import pyoracle
from mhfem_amd import problems


def _system(kind):
    pr = {"darcy3d": lambda: problems.darcy_3d(s=5, steps=1),
          "adr3d_n2": lambda: problems.adr_3d(s=3, n=2, steps=1),
          "hex_n2": lambda: problems.adr_hexes(s=3, n=2, steps=1)}[kind]()
    o = pyoracle.Oracle(pr)
    o.pre_iterate(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    rp, col, val, rhs = o.csr()
    return sp.csr_matrix((val.copy(), col.copy(), rp.copy())), rhs.copy()  # views into the oracle's memory


@pytest.mark.parametrize("kind", ["darcy3d", "adr3d_n2", "hex_n2"])
def test_ilu0_matches_c_oracle_and_reproduces_a_on_its_pattern(kind):
    A, _ = _system(kind)
    L, U = isai.ilu0(A)
    lu = pyoracle.ilu0(A.indptr, A.indices, A.data)
    LU = sp.csr_matrix((lu, A.indices, A.indptr))
    assert (L - (sp.tril(LU, -1) + sp.identity(A.shape[0]))).count_nonzero() == 0
    assert (U - sp.triu(LU)).count_nonzero() == 0
    P = (L @ U).tocsr()
    mask = sp.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
    diff = (P - A).multiply(mask)
    assert abs(diff).max() <= 1e-12 * abs(A).max()


def test_tridiagonal_closed_forms():
    rng = np.random.default_rng(4)
    m = 12
    lo, up = -rng.uniform(0.1, 1, m - 1), -rng.uniform(0.1, 1, m - 1)
    A = sp.diags([lo, 4 + rng.uniform(0, 1, m), up], [-1, 0, 1], format="csr")
    L, U = isai.ilu0(A)
    assert np.allclose((L @ U).toarray(), A.toarray(), rtol=0, atol=1e-14)  # exact LU
    ML = isai.isai(L, True).toarray()
    l = L.diagonal(-1)
    for i in range(m):
        assert ML[i, i] == 1.0
        if i >= 1:
            assert ML[i, i - 1] == pytest.approx(-l[i - 1], rel=1e-14)
        if i >= 2:
            assert ML[i, i - 2] == pytest.approx(l[i - 1] * l[i - 2], rel=1e-14)
        assert np.all(ML[i, :max(0, i - 2)] == 0)
    # the exact inverse agrees on the pattern
    Linv = sla.solve_triangular(L.toarray(), np.eye(m), lower=True)
    mask = ML != 0
    assert np.allclose(ML[mask], Linv[mask], rtol=1e-13, atol=0)


@pytest.mark.parametrize("kind", ["darcy3d", "hex_n2"])
def test_isai_rows_invert_the_factor_on_their_pattern(kind):
    A, _ = _system(kind)
    L, U = isai.ilu0(A)
    for T, lower in ((L, True), (U, False)):
        M = isai.isai(T, lower)
        P = isai.pattern_power(T, 2)
        assert np.array_equal(M.indptr, P.indptr) and np.array_equal(M.indices, P.indices)
        prod = (M @ T).tocsr()
        for i in range(0, A.shape[0], 7):
            J = M.indices[M.indptr[i]:M.indptr[i + 1]]
            row = prod[i].toarray().ravel()[J]
            e = (J == i).astype(float)
            assert np.abs(row - e).max() <= 1e-11, (lower, i)


def test_full_pattern_gives_the_exact_inverse():
    A, _ = _system("darcy3d")
    A = A[:60, :60].tocsr()
    L, U = isai.ilu0(A)
    ML, MU = isai.isai(L, True, k=60), isai.isai(U, False, k=60)
    Linv = sla.solve_triangular(L.toarray(), np.eye(60), lower=True)
    Uinv = sla.solve_triangular(U.toarray(), np.eye(60), lower=False)
    assert np.abs(ML.toarray() - Linv).max() <= 1e-11 * np.abs(Linv).max()
    assert np.abs(MU.toarray() - Uinv).max() <= 1e-11 * np.abs(Uinv).max()


def test_ilu_isai_iterations_close_to_exact_ilu():
    A, rhs = _system("darcy3d")
    x0 = np.zeros(A.shape[0])
    P = isai.IluIsai(A)
    lu = pyoracle.ilu0(A.indptr, A.indices, A.data)
    _, it_isai, res = isai.bicgstab(A, rhs, x0, P, 1e-10)
    _, it_ilu, _ = isai.bicgstab(A, rhs, x0, lambda v: pyoracle.ilu0_apply(A.indptr, A.indices, lu, v), 1e-10)
    _, it_none, _ = isai.bicgstab(A, rhs, x0, lambda v: v, 1e-10)
    assert res < 1e-10
    assert it_isai <= it_ilu + 4 and 2 * it_isai < it_none, (it_isai, it_ilu, it_none)
    # and this BiCGStab is the C oracle's (with ILU(0), where the count is not
    # sensitive to the order of the C oracle's threaded reductions)
    _, conv, it_c, _ = pyoracle.krylov(A.indptr, A.indices, A.data, rhs, x0, "bicgstab", "ilu0", 1e-10)
    assert abs(it_c - it_ilu) <= 1


def test_isai_rows_equal_the_full_construction():
    A, _ = _system("adr3d_n2")
    L, U = isai.ilu0(A)
    rows = [0, 1, 17, A.shape[0] // 2, A.shape[0] - 1]
    for T, lower in ((L, True), (U, False)):
        M = isai.isai(T, lower)
        for i, (J, v) in zip(rows, isai.isai_rows(T, rows, lower)):
            assert np.array_equal(J, M.indices[M.indptr[i]:M.indptr[i + 1]])
            assert np.array_equal(v, M.data[M.indptr[i]:M.indptr[i + 1]])


def test_parilu_fixed_point_is_ilu0():
    """ParILU's synchronous sweeps reach ILU(0) exactly once their number
    exceeds the depth of the factors' dependency chains (here a small system),
    and move towards it before that."""
    A, _ = _system("darcy3d")
    A = A[:200, :200].tocsr()
    L0, U0 = isai.ilu0(A)
    ref = abs(L0).max() + abs(U0).max()
    errs = []
    for sweeps in (0, 1, 2, 3, 5):
        L, U = isai.parilu(A, sweeps)
        errs.append(max(abs(L - L0).max(), abs(U - U0).max()) / ref)
    assert errs[0] > 1e-3 and errs[-1] < errs[0] / 10, errs
    L, U = isai.parilu(A, 200)
    assert max(abs(L - L0).max(), abs(U - U0).max()) <= 1e-13 * ref


def test_parilu_initial_guess_and_one_sweep_by_hand():
    """Zero sweeps: l_ij = a_ij / a_jj, u_ij = a_ij.  One sweep on a 3x3 matrix
    with a full pattern, worked by hand from those values."""
    a = np.array([[4.0, 1.0, 2.0], [1.0, 5.0, 1.0], [2.0, 1.0, 6.0]])
    A = sp.csr_matrix(a)
    L, U = isai.parilu(A, 0)
    assert np.allclose(L.toarray(), [[1, 0, 0], [0.25, 1, 0], [0.5, 0.2, 1]], rtol=0, atol=1e-15)
    assert np.allclose(U.toarray(), np.triu(a), rtol=0, atol=1e-15)
    L, U = isai.parilu(A, 1)
    # from l10 = .25, l20 = .5, l21 = .2, u = triu(a):
    # u11 = 5 - l10 u01 = 4.75; u12 = 1 - l10 u02 = 0.5; u22 = 6 - l20 u02 - l21 u12 = 4.8
    # l21 = (1 - l20 u01) / u11 = 0.5 / 5 = 0.1 (old u11 = 5)
    assert np.allclose(U.toarray(), [[4, 1, 2], [0, 4.75, 0.5], [0, 0, 4.8]], rtol=0, atol=1e-15)
    assert np.allclose(L.toarray(), [[1, 0, 0], [0.25, 1, 0], [0.5, 0.1, 1]], rtol=0, atol=1e-15)
