"""The AMG preconditioner on the GPU against its oracle (oracle/amg_oracle.py):
the hierarchy (level sizes, nonzeros, every aggregate) bit for bit, one
V-cycle to 1e-12, AMG-preconditioned solves against the scheme oracle, the
reuse rule over a trajectory, and the configurations it refuses."""
import numpy as np
import pytest
import scipy.sparse as sp

import amg_oracle as amg
import pyoracle
from mhfem_amd import capi, problems

pytestmark = pytest.mark.gpu

CASES = {
    "adr3d_n2": lambda: problems.adr_3d(s=4, n=2, steps=5, preconditioner="diagonal"),
    "heat2d": lambda: problems.heat_2d(s=12, steps=5),
    "hex_n3": lambda: problems.adr_hexes(s=3, n=3, steps=5),
}


def _first_solve(pr):
    s = capi.Solver(pr, preconditioner="amg", tol=1e-13)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    st = s.solveLinearSystem()
    return s, sp.csr_matrix((val, col, rp)), rhs, st


@pytest.mark.parametrize("name", sorted(CASES))
def test_hierarchy_and_vcycle_match_oracle(name):
    print(f"\n[case] {name}", flush=True)
    pr = CASES[name]()
    s, A, rhs, st = _first_solve(pr)
    levels, rebuilds = s.amg_hierarchy()
    assert rebuilds == 1
    H = amg.Hierarchy(A, pr.n)
    assert len(levels) == len(H.A)
    for l, (rows, nnz, agg) in enumerate(levels):
        assert rows == H.A[l].shape[0] and nnz == H.A[l].nnz, l
        if l + 1 < len(levels):
            assert np.array_equal(agg, H.agg[l]), l
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    got, ref = s.amg_apply(b), H.vcycle(b)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    s.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_amg_solves_match_the_scheme_oracle(name):
    print(f"\n[case] {name}", flush=True)
    pr = CASES[name]()
    s = capi.Solver(pr, preconditioner="amg", tol=1e-13)
    j = capi.Solver(pr, preconditioner="diagonal", tol=1e-13)
    o = pyoracle.Oracle(pr)
    t = 0.0
    its, its_j = [], []
    rule, rebuilds = amg.ReuseRule(), 0
    for k in range(pr.steps):
        rebuild = rule.rebuild()
        rebuilds += rebuild
        st = s.step(t, pr.dt)
        rule.after(st.iterations, rebuild)
        its.append(st.iterations)
        its_j.append(j.step(t, pr.dt).iterations)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        for nm in ("Z_iF", "Z_iK"):
            a, b = s.get_array(nm), o.array(nm)
            assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-10, (k, nm)
    print("iterations amg", its, "jacobi", its_j, flush=True)
    assert sum(its) * 2 < sum(its_j)
    # the reuse rule (Solver.hpp:1017, 1158-1160) replayed on the same iteration counts
    assert s.amg_hierarchy()[1] == rebuilds
    s.close()
    j.close()


def test_amg_refuses_fp32_and_graph_steps():
    pr = problems.heat_2d(s=6, steps=1)
    with pytest.raises(capi.MHFEMError, match="FP64"):
        capi.Solver(pr, preconditioner="amg", precision="fp32-reliable")
    s = capi.Solver(pr, preconditioner="amg")
    with pytest.raises(capi.MHFEMError, match="AMG"):
        s.step_graph(0.0, pr.dt)
    s.close()
