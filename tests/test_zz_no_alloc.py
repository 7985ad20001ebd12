"""mhfem_solve_rhs_noalloc -- mhfem_solve_rhs with its staging in the
library's fixed device array instead of per-call allocations -- against
mhfem_solve_rhs, through the C ABI: two calls in a row, the second on a
smaller system, give identical iterates and iteration counts; oversized
systems and configurations whose solve would allocate are refused."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mhfem_amd import capi, problems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def max_rows():
    text = open(os.path.join(ROOT, "include", "mhfem.h")).read()
    return int(re.search(r"#define\s+MHFEM_NOALLOC_MAX_ROWS\s+(\d+)", text).group(1))


def _assembled(pr, **kw):
    s = capi.Solver(pr, **kw)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    _, _, _, rhs = s.export_system()
    return s, rhs, s.get_array("Z_iF")


def _noalloc(s, rhs, x0):
    x = np.array(x0, np.float64, copy=True)
    b = np.ascontiguousarray(rhs, np.float64)
    st = capi.SolveStats()
    rc = capi.lib().mhfem_solve_rhs_noalloc(s.ctx, b.ctypes.data, x.ctypes.data, C.byref(st))
    return rc, x, st


@pytest.mark.parametrize("prec,precision", [("diagonal", "fp64"), ("ilu0", "fp64"), ("none", "fp64"),
                                            ("diagonal", "fp32-reliable"), ("ilu0", "fp32-refined")])
def test_noalloc_equals_solve_rhs(prec, precision):
    # two calls in a row, the second on a smaller system; each against
    # mhfem_solve_rhs on an identical context (same kernels, same inputs)
    for pr in (problems.adr_3d(s=10, steps=1, preconditioner=prec), problems.darcy_3d(s=4, steps=1, preconditioner=prec)):
        print(f"\n[noalloc] {pr.name} rows={pr.n * pr.num_faces} {prec} {precision}", flush=True)
        kw = dict(tol=1e-10, precision=precision, inner_tol=0.03, check_every=2)
        a, rhs, x0 = _assembled(pr, **kw)
        b, rhs_b, x0_b = _assembled(pr, **kw)
        assert np.array_equal(rhs, rhs_b) and np.array_equal(x0, x0_b)
        x_ref, st_ref = a.solve_rhs(rhs, x0)
        rc, x, st = _noalloc(b, rhs, x0)
        assert rc == 0, capi.lib().mhfem_last_error(b.ctx)
        assert st.converged and st.iterations == st_ref.iterations
        assert st.residue == st_ref.residue
        assert np.array_equal(x, x_ref)
        # a second call on the same context from the same start repeats it
        rc, x2, st2 = _noalloc(b, rhs, x0)
        assert rc == 0 and st2.iterations == st.iterations and np.array_equal(x2, x)
        a.close()
        b.close()


def test_noalloc_refusals():
    pr = problems.heat_2d(s=6, steps=1)
    for kw, msg in ((dict(solver="gmres"), b"BiCGStab"), (dict(preconditioner="amg"), b"BiCGStab"),
                    (dict(preconditioner="ilu-isai"), b"BiCGStab")):
        s, rhs, x0 = _assembled(pr, **kw)
        rc, _, _ = _noalloc(s, rhs, x0)
        assert rc == -1
        assert msg in capi.lib().mhfem_last_error(s.ctx)
        s.close()
    s, rhs, x0 = _assembled(pr)
    s.configure_guess(True, train_steps=0, history_depth=2)
    rc, _, _ = _noalloc(s, rhs, x0)
    assert rc == -1 and b"learned guess" in capi.lib().mhfem_last_error(s.ctx)
    s.close()
    # a system one size past the staging array
    big = problems.adr_3d(s=36, n=2, steps=1, preconditioner="diagonal")
    assert big.n * big.num_faces > max_rows()
    s = capi.Solver(big)
    rows = big.n * big.num_faces
    rc, _, _ = _noalloc(s, np.zeros(rows), np.zeros(rows))
    assert rc == -1 and b"at most" in capi.lib().mhfem_last_error(s.ctx)
    s.close()
