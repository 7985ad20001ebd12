"""The ILU-ISAI and ParILU-ISAI preconditioners on the GPU against their
oracle (oracle/isai_oracle.py): the approximate inverses (patterns bit for
bit, values to 1e-12), the BiCGStab iteration count from the same x0, whole
time steps against the scheme oracle, the reuse rule over a trajectory, and
the configurations they refuse."""
import numpy as np
import pytest
import scipy.sparse as sp

import amg_oracle as amg
import isai_oracle as isai
import pyoracle
from mhfem_amd import capi, problems

pytestmark = pytest.mark.gpu

CASES = {
    "darcy3d": lambda: problems.darcy_3d(s=5, steps=4),
    "adr3d_n2": lambda: problems.adr_3d(s=4, n=2, steps=4),
    "heat2d": lambda: problems.heat_2d(s=12, steps=4),
    "hex_n2": lambda: problems.adr_hexes(s=3, n=2, steps=4),
}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


FACTORS = {"ilu-isai": None, "parilu-isai": isai.PARILU_SWEEPS}


@pytest.mark.parametrize("prec", sorted(FACTORS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_inverses_and_iterations_match_oracle(name, prec):
    print(f"\n[case] {name} {prec}", flush=True)
    pr = CASES[name]()
    s = capi.Solver(pr, preconditioner=prec, tol=1e-10)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    x0 = s.get_array("Z_iF")
    x, st = s.solve_rhs(rhs, x0)
    A = sp.csr_matrix((val, col, rp))
    P = isai.IluIsai(A, sweeps=FACTORS[prec])
    (lrp, lcol, lval), (urp, ucol, uval), rebuilds = s.isai_matrices()
    assert rebuilds == 1
    for (grp, gcol, gval), M in (((lrp, lcol, lval), P.ML), ((urp, ucol, uval), P.MU)):
        assert np.array_equal(grp, M.indptr) and np.array_equal(gcol, M.indices)
        assert np.abs(gval - M.data).max() <= 1e-12 * np.abs(M.data).max()
    xo, it, res = isai.bicgstab(A, rhs, x0, P, 1e-10)
    print("iterations", st.iterations, it, flush=True)
    assert st.converged and abs(st.iterations - it) <= 2, (st.iterations, it)
    assert rel(x, xo) < 1e-7
    s.close()


@pytest.mark.parametrize("prec", sorted(FACTORS))
@pytest.mark.parametrize("name", ["darcy3d", "hex_n2"])
def test_steps_match_scheme_oracle_and_reuse_rule(name, prec):
    print(f"\n[case] {name} {prec}", flush=True)
    pr = CASES[name]()
    s = capi.Solver(pr, preconditioner=prec, tol=1e-13)
    o = pyoracle.Oracle(pr)
    rule, rebuilds, its = amg.ReuseRule(), 0, []
    t = 0.0
    for k in range(pr.steps):
        rebuild = rule.rebuild()
        rebuilds += rebuild
        st = s.step(t, pr.dt)
        rule.after(st.iterations, rebuild)
        its.append(st.iterations)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        assert st.converged
        for nm in ("Z_iF", "Z_iK"):
            assert rel(s.get_array(nm), o.array(nm)) < 1e-10, (k, nm)
    print("iterations", its, "rebuilds", rebuilds, flush=True)
    assert s.isai_matrices()[2] == rebuilds
    s.close()


def test_ilu_isai_refuses_fp32_and_graph_steps():
    pr = problems.heat_2d(s=6, steps=1)
    with pytest.raises(capi.MHFEMError, match="FP64"):
        capi.Solver(pr, preconditioner="ilu-isai", precision="fp32-reliable")
    s = capi.Solver(pr, preconditioner="ilu-isai")
    with pytest.raises(capi.MHFEMError, match="ILU-ISAI"):
        s.step_graph(0.0, pr.dt)
    s.close()


def test_large_system_sampled_rows_and_solution():
    """A system of more rows than one launch of the library's grid-capped
    kernels covers (n = 2 on 32^3 x 6 tetrahedra, about 800 k rows): sampled
    rows of M_L and M_U, the last ones included, equal the oracle's rows built
    from the C oracle's ILU(0) of the exported matrix, and the solution equals
    a plain BiCGStab solve to 1e-6 (a preconditioner with missing rows can
    still drive the preconditioned residual down)."""
    print("\n[case] adr3d n=2 32^3", flush=True)
    pr = problems.adr_3d(s=32, n=2, steps=1)
    s = capi.Solver(pr, preconditioner="ilu-isai", tol=1e-10)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    x0 = s.get_array("Z_iF")
    x, st = s.solve_rhs(rhs, x0)
    assert st.converged
    n = rp.size - 1
    assert n > 2048 * 256
    lu = pyoracle.ilu0(rp, col, val)
    LU = sp.csr_matrix((lu, col, rp), shape=(n, n))
    L = (sp.tril(LU, -1) + sp.identity(n)).tocsr()
    U = sp.triu(LU).tocsr()
    L.sort_indices()
    U.sort_indices()
    rows = [0, 1, n // 3, 2048 * 256 - 1, 2048 * 256, n - 2, n - 1]
    rows += list(np.random.default_rng(0).integers(0, n, 8))
    (lrp, lcol, lval), (urp, ucol, uval), _ = s.isai_matrices()
    for (grp, gcol, gval), T, lower in (((lrp, lcol, lval), L, True), ((urp, ucol, uval), U, False)):
        for i, (J, v) in zip(rows, isai.isai_rows(T, rows, lower)):
            assert np.array_equal(gcol[grp[i]:grp[i + 1]], J), (lower, i)
            assert np.abs(gval[grp[i]:grp[i + 1]] - v).max() <= 1e-12 * np.abs(v).max(), (lower, i)
    s.close()
    ref = capi.Solver(pr, preconditioner="diagonal", tol=1e-12)
    ref.preIterate(0.0, pr.dt)
    ref.assembleLinearSystem(0.0, pr.dt)
    xr, st_r = ref.solve_rhs(rhs, x0)
    assert st_r.converged and rel(x, xr) < 1e-6
    ref.close()
