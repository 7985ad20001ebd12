"""Batched ensembles, graph-captured steps and the autotuner on the GPU."""
import numpy as np
import pytest

from mhfem_amd import capi, problems
import pyoracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _members(sizes, seed=5):
    return [problems.ensemble_member(s, seed * 100 + k, steps=3) for k, s in enumerate(sizes)]


def test_batch_matches_individual_runs():
    """Each instance of a batch evolves as it would alone (oracle runs)."""
    members = _members([5, 9, 3, 12, 7])
    batch = problems.concatenate(members)
    s = capi.Solver(batch, tol=1e-13, check_every=2)
    co, fo = s.batch_info()
    assert np.array_equal(co, batch.extra["cell_offsets"])
    assert np.array_equal(fo, batch.extra["face_offsets"])
    oracles = [pyoracle.Oracle(m) for m in members]
    t = 0.0
    for step in range(3):
        s.step_graph(t, batch.dt)
        for o in oracles:
            o.step(t, batch.dt, tol=1e-13)
        t += batch.dt
        zf = s.get_array("Z_iF")
        zk = s.get_array("Z_iK")
        for k, o in enumerate(oracles):
            assert rel(zf[fo[k]:fo[k + 1]], o.array("Z_iF")) < 1e-10, (step, k)
            assert rel(zk[co[k]:co[k + 1]], o.array("Z_iK")) < 1e-10, (step, k)


def test_batch_iterations_per_instance():
    """Per-instance convergence: each instance stops at its own iteration
    count, as the same BiCGStab run alone (oracle) does."""
    members = _members([4, 10, 6])
    batch = problems.concatenate(members)
    s = capi.Solver(batch, tol=1e-10, check_every=1)
    s.step_graph(0.0, batch.dt)
    its, res = s.batch_stats()
    assert np.all(res < 1e-10)
    for k, m in enumerate(members):
        o = pyoracle.Oracle(m)
        it, _ = o.step(0.0, m.dt, tol=1e-10)
        assert abs(int(its[k]) - it) <= 2, (k, its[k], it)
    assert len(set(its.tolist())) > 1  # instances really stop independently


@pytest.mark.parametrize("check_every,rows_per_block", [(1, 128), (3, 0)])
def test_batch_narrowed_steps_equal_single_runs(check_every, rows_per_block):
    """Once some instances of a batch have converged, its steps launch only
    the blocks of the instances still iterating (the graph step replays its
    graph until then and iterates eagerly over the narrowed grids after).
    Skipping the halted instances' blocks changes no arithmetic: a segment's
    work-groups start at its first row, so its products and dot products are
    summed as in a context of its own.  Each instance's iteration count and
    solution equal its own single-instance run bit for bit, for the eager
    and the graph step, step after step (128 rows per work-group: several
    work-groups per instance; 0: the default, one per instance here)."""
    members = _members([4, 12, 6, 16, 3, 9, 5])
    batch = problems.concatenate(members)
    kw = dict(tol=1e-12, check_every=check_every, rows_per_block=rows_per_block)
    a = capi.Solver(batch, **kw)
    b = capi.Solver(batch, **kw)
    singles = [capi.Solver(m, **kw) for m in members]
    co, fo = a.batch_info()
    t = 0.0
    for k in range(3):
        print(f"[batch] narrowed step {k}", flush=True)
        a.step(t, batch.dt)
        b.step_graph(t, batch.dt)
        its1 = np.array([s.step(t, batch.dt).iterations for s in singles])
        t += batch.dt
        ia, _ = a.batch_stats()
        ib, _ = b.batch_stats()
        assert len(set(its1.tolist())) > 1, its1  # instances stop at different counts
        assert np.array_equal(ia, its1), (k, ia, its1)
        assert np.array_equal(ib, its1), (k, ib, its1)
        za, zb = a.get_array("Z_iF"), b.get_array("Z_iF")
        ka, kb = a.get_array("Z_iK"), b.get_array("Z_iK")
        for j, s in enumerate(singles):
            z1, k1 = s.get_array("Z_iF"), s.get_array("Z_iK")
            assert np.array_equal(za[fo[j]:fo[j + 1]], z1), (k, j)
            assert np.array_equal(zb[fo[j]:fo[j + 1]], z1), (k, j)
            assert np.array_equal(ka[co[j]:co[j + 1]], k1), (k, j)
            assert np.array_equal(kb[co[j]:co[j + 1]], k1), (k, j)


@pytest.mark.parametrize("prec", ["diagonal", "ilu0"])
def test_graph_step_equals_eager_step(prec):
    """Replaying the captured graphs gives exactly the eager result, step after
    step (eager work runs between the replays)."""
    pr = problems.darcy_3d(s=5, steps=4, preconditioner=prec)
    a = capi.Solver(pr, tol=1e-10, check_every=4)
    b = capi.Solver(pr, tol=1e-10, check_every=4)
    t = 0.0
    for k in range(4):
        sa = a.step(t, pr.dt)
        sb = b.step_graph(t, pr.dt)
        t += pr.dt
        assert sa.iterations == sb.iterations
        assert np.array_equal(a.get_array("Z_iF"), b.get_array("Z_iF"))
        assert np.array_equal(a.get_array("Z_iK"), b.get_array("Z_iK"))


def test_graph_step_after_coefficient_updates():
    """Coefficients set between graph steps -- a uniform value replaced by
    another, a uniform coefficient made a full array and back, a full array
    rewritten -- give exactly the eager steps: the graphs are captured again
    whenever the scheme's pointers or uniform masks change, and kept
    allocations stay valid for the graphs that still use them."""
    pr = problems.darcy_3d(s=5, steps=6, preconditioner="diagonal")
    a = capi.Solver(pr, tol=1e-10, check_every=4)
    b = capi.Solver(pr, tol=1e-10, check_every=4)
    rng = np.random.default_rng(2)
    C = pr.num_cells
    updates = [None, ("N_ijK", np.array([0.5])), ("N_ijK", rng.uniform(0.1, 1.0, C)),
               ("D_ijK", rng.uniform(0.5, 2.0, C)), ("N_ijK", np.array([0.3])), ("m_iK", rng.uniform(0.5, 2.0, C))]
    t = 0.0
    for k, up in enumerate(updates):
        print(f"[graph] step {k} after {up[0] if up else 'nothing'}", flush=True)
        if up:
            a.set_array(*up)
            b.set_array(*up)
        sa = a.step(t, pr.dt)
        sb = b.step_graph(t, pr.dt)
        t += pr.dt
        assert sa.iterations == sb.iterations, k
        assert np.array_equal(a.get_array("Z_iF"), b.get_array("Z_iF")), k
        assert np.array_equal(a.get_array("Z_iK"), b.get_array("Z_iK")), k


def test_autotune_keeps_results():
    pr = problems.darcy_3d(s=8, steps=2, preconditioner="diagonal")
    a = capi.Solver(pr, tol=1e-13)
    b = capi.Solver(pr, tol=1e-13)
    rpb, ms = b.autotune()
    assert rpb in (256, 512, 1024, 2048, 4096, 8192) and ms > 0
    t = 0.0
    for k in range(2):
        a.step(t, pr.dt)
        b.step(t, pr.dt)
        t += pr.dt
    # a different work decomposition changes only the summation order
    assert rel(b.get_array("Z_iF"), a.get_array("Z_iF")) < 1e-10
