"""Time-dependent boundary models evaluated on the device (the reference's
BoundaryModel argument, Solver.h:28-31, called at time + tau inside the
assembly: Solver.hpp:948, BoundaryConditions.hpp:379, 414, 455), against the
oracle, which evaluates the same model on the host; and the learned-guess
workload on the time-dependent n = 2 Kuhn system."""
import dataclasses

import numpy as np
import pytest

from mhfem_amd import capi, problems
import pyoracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _sine_problem(s=4, dt=0.1, steps=6):
    """Z_1 = sin(2 pi t / 0.5) at x = 0; with dt = 0.1 the Dirichlet value
    takes both signs and passes through zero over six steps."""
    return problems.adr_3d(s=s, dt=dt, steps=steps, preconditioner="diagonal", time_dependent=True)


def test_time_dependent_dirichlet_against_oracle():
    pr = _sine_problem()
    s = capi.Solver(pr, tol=1e-13)
    o = pyoracle.Oracle(pr)
    t = 0.0
    x0 = np.nonzero(pr.bc_dirichlet[:pr.num_faces] == 1.0)[0]
    assert x0.size > 0
    for k in range(pr.steps):
        print(f"[boundary model] sine, step {k}", flush=True)
        s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        # the right-hand side of a Dirichlet row is the model's value at t + tau
        # (the last assembly's; the stored values stay untouched)
        rhs = s.get_array("rhs")
        want = np.sin(2 * np.pi * t / 0.5)
        assert np.allclose(rhs[x0 * pr.n], want, rtol=0, atol=1e-15), k
        assert np.array_equal(s.get_array("bc_dirichlet"), pr.bc_dirichlet)
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10, k
        assert rel(s.get_array("Z_iK"), o.array("Z_iK")) < 1e-10, k


def test_table_and_neumann_time_functions():
    """Every kind of time function, on Dirichlet and Neumann values of a
    problem with all four boundary types: a table (linear in between,
    constant outside), a sine with offset and phase, a constant."""
    pr = problems.adr_3d(s=4, dt=0.05, steps=5, preconditioner="diagonal")
    rng = np.random.default_rng(3)
    n, F = pr.n, pr.num_faces
    tags = np.array(pr.tags, copy=True)
    vals = np.zeros(n * F)
    dirv = np.zeros(n * F)
    kinds = [problems.FIXED_VALUE, problems.FIXED_FLUX, problems.ADVECTIVE_OUTFLOW,
             problems.FIXED_FLUX_NO_ADVECTION]
    for i in range(n):
        for q, E in enumerate(np.nonzero(pr.face_cells[:, 1] < 0)[0]):
            tags[i * F + E] = kinds[(q + i) % 4]
            vals[i * F + E] = rng.uniform(-1, 1)
            dirv[i * F + E] = rng.uniform(-1, 1)
    funcs = {("dirichlet", 0): dict(kind="table", times=[0.04, 0.1, 0.13, 0.2], values=[1.0, -0.5, 2.0, 0.25]),
             ("dirichlet", 1): dict(kind="sine", offset=0.3, amplitude=0.7, angular_frequency=9.0, phase=0.4),
             ("neumann", 0): dict(kind="constant", offset=-1.5),
             ("neumann", 1): dict(kind="table", times=[0.0, 0.3], values=[0.0, 3.0])}
    pr = dataclasses.replace(pr, tags=tags, bc_values=vals, bc_dirichlet=dirv, boundary_functions=funcs)
    s = capi.Solver(pr, tol=1e-13)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        print(f"[boundary model] table/sine/constant, step {k}", flush=True)
        s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        assert rel(s.get_array("rhs"), o.array("rhs")) < 1e-12, k
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10, k
    # back to the stationary model for species 0's Dirichlet values
    s.set_boundary_time_function("dirichlet", 0, None)
    pr2 = dataclasses.replace(pr, boundary_functions={kk: v for kk, v in funcs.items() if kk != ("dirichlet", 0)})
    o.problem = pr2
    s.step(t, pr.dt)
    o.step(t, pr.dt, tol=1e-13)
    assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10


def test_boundary_time_function_refusals():
    pr = problems.adr_3d(s=2, steps=1, preconditioner="diagonal")
    s = capi.Solver(pr)
    with pytest.raises(capi.MHFEMError, match="ascending"):
        s.set_boundary_time_function("dirichlet", 0, dict(kind="table", times=[0.0, 0.0], values=[1.0, 2.0]))
    with pytest.raises(capi.MHFEMError, match="species"):
        s.set_boundary_time_function("neumann", 2, dict(kind="constant", offset=1.0))
    L = capi.lib()
    assert L.mhfem_set_boundary_time_function(s.ctx, capi.ARRAY_ID["Z_iF"], 0, None) == -1
    s.close()


def test_graph_steps_see_each_steps_time():
    """The captured assembly reads the model's time from device memory: graph
    replays equal eager steps bit for bit while the Dirichlet value moves."""
    pr = _sine_problem(s=4, dt=0.1, steps=5)
    a = capi.Solver(pr, tol=1e-10, check_every=4)
    b = capi.Solver(pr, tol=1e-10, check_every=4)
    t = 0.0
    for k in range(pr.steps):
        print(f"[boundary model] graph replay, step {k}", flush=True)
        sa = a.step(t, pr.dt)
        sb = b.step_graph(t, pr.dt)
        t += pr.dt
        assert sa.iterations == sb.iterations, k
        assert np.array_equal(a.get_array("Z_iF"), b.get_array("Z_iF")), k
        assert np.array_equal(a.get_array("rhs"), b.get_array("rhs")), k


def test_learned_guess_on_the_time_dependent_system():
    """The learned-guess workload (the n = 2 Kuhn system with Z_1 =
    sin(2 pi t / 0.5) at x = 0) on a small mesh: the network is trained
    online on the first steps, then the FP32-reliable form with the learned
    starting vector stays within 1e-10 of the all-FP64 oracle at every step."""
    pr = problems.adr_3d(s=5, dt=0.02, steps=8, preconditioner="diagonal", time_dependent=True)
    s = capi.Solver(pr, tol=1e-13, precision="fp32-reliable", inner_tol=0.03)
    s.configure_guess(True, train_steps=4, epochs=4, learning_rate=2e-3, sample_stride=1, history_depth=4)
    o = pyoracle.Oracle(pr)
    t = 0.0
    used = 0
    for k in range(pr.steps):
        print(f"[boundary model] learned guess, step {k}", flush=True)
        st = s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        g = s.guess_stats()
        used += g["used"]
        assert st.converged and st.residue < 1e-13
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10, k
        assert rel(s.get_array("Z_iK"), o.array("Z_iK")) < 1e-10, k
    assert used > 0
    assert s.guess_stats()["trained_steps"] == 4
