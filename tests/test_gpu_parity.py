"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: integer structure (face numbering, incidence, CSR rowptr/colidx)
bit-exact; FP64 scheme arrays and matrix values within 1e-12 relative (only
summation order and FMA contraction differ); face traces Z_iF and cell
averages Z_iK within 1e-10 relative L2 after each time step when both sides
solve to 1e-13 (solver-independent comparison, tolerances in each test).
"""
import dataclasses

import numpy as np
import pytest

from mhfem_amd import meshgen, problems, capi
import pyoracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def maxrel(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s


def _mixed_tags(pr, seed=5):
    """Every boundary face cycles through the four tags (per species, offset
    by the species), with random flux and Dirichlet values."""
    rng = np.random.default_rng(seed)
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
    return dataclasses.replace(pr, tags=tags, bc_values=vals, bc_dirichlet=dirv)


def _uneven_boxes(pr, seed=7):
    """The same box mesh with uneven, anisotropic spacing (h_x != h_y): the
    widths along each axis vary by up to 3x, so the systems stay as well
    conditioned as the uniform ones and the 1e-10 bar measures the method."""
    rng = np.random.default_rng(seed)
    pts = np.array(pr.points, copy=True)
    for a in range(pr.dim):
        lines = np.unique(pts[:, a])
        w = rng.uniform(0.5, 1.5, lines.size - 1)
        moved = np.concatenate([[0.0], np.cumsum(w) / w.sum()]) * (1.0 - 0.3 * a)
        pts[:, a] = moved[np.searchsorted(lines, pts[:, a])]
    return dataclasses.replace(pr, points=pts)


def _mobility(pr, seed=9):
    """The upwinded-mobility model (do_mobility_upwind) with a random,
    non-uniform m_iK and a boundary mobility different from it, so that the
    choice of upwind cell changes the values.  Boundary faces get FixedValue,
    FixedFlux or FixedFluxNoAdvection with fluxes away from zero for every
    species: a zero flux (AdvectiveOutflow, a no-flow wall) leaves the
    any-species inflow test (Solver.hpp:680-690) to rounding."""
    rng = np.random.default_rng(seed)
    n, C, F = pr.n, pr.num_cells, pr.num_faces
    tags = np.array(pr.tags, copy=True)
    vals = np.zeros(n * F)
    dirv = np.zeros(n * F)
    kinds = [problems.FIXED_VALUE, problems.FIXED_FLUX, problems.FIXED_FLUX_NO_ADVECTION]
    for i in range(n):
        for q, E in enumerate(np.nonzero(pr.face_cells[:, 1] < 0)[0]):
            tags[i * F + E] = kinds[(q + i) % 3]
            vals[i * F + E] = rng.choice([-1.0, 1.0]) * rng.uniform(0.2, 1.0)
            dirv[i * F + E] = rng.uniform(-1, 1)
    co = dict(pr.coeffs, m_iK=rng.uniform(0.5, 2.0, n * C))
    return dataclasses.replace(pr, coeffs=co, tags=tags, bc_values=vals, bc_dirichlet=dirv, do_mobility_upwind=1,
                               bc_mobility=rng.uniform(0.5, 2.0, n * F), Z_iK0=rng.uniform(0, 1, n * C))


CASES = {
    "heat2d": lambda: problems.heat_2d(s=12, steps=3),
    "heat2d_lumped": lambda: problems.heat_2d(s=12, steps=3, lumping=1),
    "darcy3d": lambda: problems.darcy_3d(s=6, steps=3, preconditioner="diagonal"),
    "darcy3d_lumped": lambda: dataclasses.replace(problems.darcy_3d(s=5, steps=3, preconditioner="diagonal"), lumping=1),
    "adr3d_n2": lambda: problems.adr_3d(s=5, steps=3, preconditioner="diagonal"),
    "adr3d_implicit_upwind": lambda: dataclasses.replace(problems.adr_3d(s=4, steps=2, preconditioner="diagonal"),
                                                         advection=problems.IMPLICIT_UPWIND),
    "adr3d_implicit_trace": lambda: dataclasses.replace(problems.adr_3d(s=4, steps=2, preconditioner="diagonal"),
                                                        advection=problems.IMPLICIT_TRACE),
    "adr3d_n3": lambda: problems.adr_3d(s=3, steps=2, preconditioner="diagonal", n=3),
    "adr3d_n3_implicit_upwind": lambda: dataclasses.replace(
        problems.adr_3d(s=3, steps=2, preconditioner="diagonal", n=3), advection=problems.IMPLICIT_UPWIND),
    "adr2d_n2": lambda: problems.adr_2d(s=8, n=2),
    "adr3d_mixed_tags": lambda: _mixed_tags(problems.adr_3d(s=4, steps=2, preconditioner="diagonal")),
    "adr3d_n3_mixed_tags_upwind": lambda: _mixed_tags(dataclasses.replace(
        problems.adr_3d(s=3, steps=2, preconditioner="diagonal", n=3), advection=problems.IMPLICIT_UPWIND)),
    "adr2d_n3_implicit_trace": lambda: dataclasses.replace(problems.adr_2d(s=6, n=3),
                                                           advection=problems.IMPLICIT_TRACE),
    # box cells: quadrangles (axis-aligned rectangles) and edges (MassMatrix.h:28-303)
    "quad_heat": lambda: problems.heat_quads(s=10, steps=3),
    "quad_heat_lumped": lambda: _uneven_boxes(problems.heat_quads(s=10, steps=3, lumping=1)),
    "quad_adr_n2": lambda: _uneven_boxes(problems.adr_quads(s=8, n=2)),
    "quad_adr_n2_implicit_upwind": lambda: dataclasses.replace(_uneven_boxes(problems.adr_quads(s=6, n=2)),
                                                               advection=problems.IMPLICIT_UPWIND),
    "quad_adr_n3_lumped_mixed_tags": lambda: _mixed_tags(_uneven_boxes(problems.adr_quads(s=6, n=3, lumping=1))),
    "edge_adr_n1": lambda: _uneven_boxes(problems.adr_1d(s=40, n=1)),
    "edge_adr_n2_lumped": lambda: _uneven_boxes(problems.adr_1d(s=40, n=2, lumping=1)),
    "edge_adr_n3_mixed_tags_trace": lambda: _mixed_tags(dataclasses.replace(
        _uneven_boxes(problems.adr_1d(s=30, n=3)), advection=problems.IMPLICIT_TRACE)),
    "hex_adr_n1_lumped": lambda: _uneven_boxes(problems.adr_hexes(s=4, n=1, lumping=1)),
    "hex_adr_n2": lambda: _uneven_boxes(problems.adr_hexes(s=4, n=2)),
    "hex_adr_n2_implicit_upwind": lambda: dataclasses.replace(_uneven_boxes(problems.adr_hexes(s=3, n=2)),
                                                              advection=problems.IMPLICIT_UPWIND),
    "hex_adr_n3_mixed_tags": lambda: _mixed_tags(_uneven_boxes(problems.adr_hexes(s=3, n=3))),
    # non-uniform mobility with mobility upwinding (Solver.hpp:663-727)
    "adr3d_n2_mobility": lambda: _mobility(problems.adr_3d(s=4, steps=3, preconditioner="diagonal")),
    "darcy3d_mobility": lambda: _mobility(problems.darcy_3d(s=5, steps=3, preconditioner="diagonal")),
    "adr2d_n3_mobility": lambda: _mobility(problems.adr_2d(s=6, n=3, steps=3)),
    "hex_adr_n2_mobility": lambda: _mobility(_uneven_boxes(problems.adr_hexes(s=3, n=2, steps=3))),
    "quad_adr_n2_mobility_implicit_upwind": lambda: dataclasses.replace(
        _mobility(_uneven_boxes(problems.adr_quads(s=6, n=2, steps=3))), advection=problems.IMPLICIT_UPWIND),
}


@pytest.fixture(params=sorted(CASES))
def case(request):
    print(f"\n[case] {request.param}", flush=True)
    return request.param, CASES[request.param]()


def test_topology_bit_exact(case):
    name, pr = case
    s = capi.Solver(pr, setup=False)
    cf, fc = s.topology()
    ocf, ofc = pyoracle.topology(pr.points, pr.cells)
    assert np.array_equal(cf, ocf)
    assert np.array_equal(fc, ofc)
    assert np.array_equal(cf, pr.cell_faces)
    s.close()


def test_measures(case):
    name, pr = case
    s = capi.Solver(pr, setup=False)
    o = pyoracle.Oracle(pr)
    assert maxrel(s.get_array("cell_measure"), o.array("cell_measure")) < 1e-14
    assert maxrel(s.get_array("face_measure"), o.array("face_measure")) < 1e-14


def test_pre_iterate_arrays(case):
    name, pr = case
    s = capi.Solver(pr, tol=1e-13)
    o = pyoracle.Oracle(pr)
    assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-15
    s.preIterate(0.0, pr.dt)
    o.pre_iterate(0.0, pr.dt)
    for nm in ("b_ijK_storage", "R_ijKe", "R_iK", "m_iE_upw", "Z_ijE_upw"):
        a, b = s.get_array(nm), o.array(nm)
        assert maxrel(a, b) < 1e-12, (nm, maxrel(a, b))


def test_assembly(case):
    name, pr = case
    s = capi.Solver(pr, tol=1e-13)
    o = pyoracle.Oracle(pr)
    s.preIterate(0.0, pr.dt)
    o.pre_iterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    orp, ocol, oval, orhs = o.csr()
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert maxrel(val, oval) < 1e-12
    assert maxrel(rhs, orhs) < 1e-12
    # the diagonal of every row is exactly 1 (LinearSystem.h:196-201, BoundaryConditions.hpp:200-203, :375-377)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert np.all(val[col == rows] == 1.0)


def test_time_steps(case):
    name, pr = case
    s = capi.Solver(pr, tol=1e-13)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        eF = rel(s.get_array("Z_iF"), o.array("Z_iF"))
        eK = rel(s.get_array("Z_iK"), o.array("Z_iK"))
        assert eF < 1e-10 and eK < 1e-10, (k, eF, eK)
        if pr.do_mobility_upwind:
            assert rel(s.get_array("v_iKe"), o.array("v_iKe")) < 1e-9
            # the upwind choices of this step's preIterate: copies of stored
            # mobilities or means of two, so identical
            assert np.array_equal(s.get_array("m_iE_upw"), o.array("m_iE_upw")), k


def test_mobility_upwind_of_assigned_fluxes():
    """K4 on the GPU from fluxes set directly (the oracle's pin of the same
    rule): species 0 and 1 carry the face fluxes of two constant fields of
    different directions, species 2 none; the HIP k_upwind and the oracle
    take identical decisions on every face."""
    pr = _mobility(problems.adr_3d(s=3, steps=1, preconditioner="diagonal", n=3))
    n, C, F = pr.n, pr.num_cells, pr.num_faces
    fv = meshgen.outward_face_vectors(pr.points, pr.cells)  # (C, fpc, 3): |E| n
    u = np.array([[1.0, 0.37, 0.21], [-0.3, 0.8, -0.45], [0.0, 0.0, 0.0]])
    v = np.einsum("id,ked->iek", u, fv).ravel()
    s = capi.Solver(pr, tol=1e-13)
    o = pyoracle.Oracle(pr)
    s.set_array("v_iKe", v)
    o.array("v_iKe")[:] = v
    s.preIterate(0.0, pr.dt)
    o.pre_iterate(0.0, pr.dt)
    got, ref = s.get_array("m_iE_upw"), o.array("m_iE_upw")
    assert np.array_equal(got, ref)
    # the data exercises both upwind cells, the tie and boundary inflow
    m = pr.coefficient_array("m_iK").reshape(n, C)
    inner = pr.face_cells[:, 1] >= 0
    K1, K2 = pr.face_cells[inner, 0], pr.face_cells[inner, 1]
    g = got.reshape(F, n)[inner]
    assert np.any(g[:, 0] == m[0, K1]) and np.any(g[:, 0] == m[0, K2])
    assert np.allclose(g[:, 2], 0.5 * (m[2, K1] + m[2, K2]))
    bnd = ~inner
    assert np.any(got.reshape(F, n)[bnd] == pr.bc_mobility.reshape(n, F).T[bnd])
    s.close()


@pytest.mark.parametrize("solver,prec", [("bicgstab", "diagonal"), ("bicgstab", "none"), ("bicgstab", "ilu0"),
                                         ("gmres", "diagonal"), ("gmres", "ilu0"), ("bicgstab", "ilut"),
                                         ("gmres", "ilut"), ("tfqmr", "ilut")])
def test_krylov_iteration_parity(solver, prec):
    """Same algorithm, same x0: iteration counts agree and the iterates match."""
    print(f"\n[krylov] {solver} {prec}", flush=True)
    pr = problems.darcy_3d(s=6, steps=1, preconditioner=prec)
    s = capi.Solver(pr, solver=solver, preconditioner=prec, tol=1e-10, check_every=1)
    o = pyoracle.Oracle(pr)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    x0 = s.get_array("Z_iF")
    x, st = s.solve_rhs(rhs, x0)
    xo, conv, it, res = pyoracle.krylov(rp, col, val, rhs, x0, solver, prec, 1e-10)
    assert conv == 1 and st.converged
    assert abs(st.iterations - it) <= 2, (st.iterations, it)
    assert rel(x, xo) < 1e-7
    # both are solutions to the requested residue
    xd, _, _, _ = pyoracle.krylov(rp, col, val, rhs, x0, "bicgstab", "ilu0", 1e-14)
    assert rel(x, xd) < 1e-6


_SYSTEMS = {"heat2d": lambda prec: problems.heat_2d(s=12, steps=1),
            "darcy3d": lambda prec: problems.darcy_3d(s=6, steps=1, preconditioner=prec)}


@pytest.mark.parametrize("solver,prec,case", [
    ("cg", "none", "heat2d"), ("cg", "diagonal", "heat2d"), ("cg", "ilu0", "darcy3d"),
    ("tfqmr", "none", "darcy3d"), ("tfqmr", "diagonal", "darcy3d"), ("tfqmr", "ilu0", "darcy3d"),
    ("tfqmr", "ilu0", "heat2d"), ("jacobi", "none", "heat2d"), ("jacobi", "ilu0", "darcy3d")])
def test_other_solvers_iteration_parity(solver, prec, case):
    """CG, TFQMR (with its restarts) and the Jacobi iteration against the
    oracle's from the same x0: iteration counts agree and so do the iterates.
    Unpreconditioned TFQMR on the Darcy system is sensitive to rounding: the
    oracle's own count moves between 79 and 95 when the matrix entries are
    perturbed by one or two ulps, so TFQMR's counts are compared within 20 %;
    the others within 2 %."""
    print(f"\n[krylov] {solver} {prec} {case}", flush=True)
    pr = _SYSTEMS[case](prec)
    s = capi.Solver(pr, solver=solver, preconditioner=prec, tol=1e-10)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    x0 = s.get_array("Z_iF")
    x, st = s.solve_rhs(rhs, x0)
    xo, conv, it, res = pyoracle.krylov(rp, col, val, rhs, x0, solver, prec, 1e-10, max_iter=20000)
    assert conv == 1 and st.converged and st.residue < 1e-10
    assert abs(st.iterations - it) <= max(2, (0.2 if solver == "tfqmr" else 0.02) * it), (st.iterations, it)
    assert rel(x, xo) < 1e-7
    xd, _, _, _ = pyoracle.krylov(rp, col, val, rhs, x0, "bicgstab", "ilu0", 1e-14)
    assert rel(x, xd) < 1e-6
    s.close()


@pytest.mark.parametrize("solver,m,prec,case", [
    ("bicgstabl", 1, "none", "darcy3d"), ("bicgstabl", 2, "diagonal", "darcy3d"), ("bicgstabl", 4, "ilu0", "darcy3d"),
    ("bicgstabl", 2, "none", "heat2d"), ("idrs", 1, "ilu0", "darcy3d"), ("idrs", 4, "none", "darcy3d"),
    ("idrs", 4, "ilu0", "darcy3d"), ("idrs", 8, "diagonal", "heat2d")])
def test_bicgstabl_and_idrs_iteration_parity(solver, m, prec, case):
    """BiCGStab(l) and IDR(s) (same shadow space, from the same integer hash)
    against the oracle's from the same x0: iteration counts within 10 % and
    the iterates within 1e-7."""
    print(f"\n[krylov] {solver}({m}) {prec} {case}", flush=True)
    pr = _SYSTEMS[case](prec)
    kw = {"bicgstab_ell": m} if solver == "bicgstabl" else {"idr_s": m}
    s = capi.Solver(pr, solver=solver, preconditioner=prec, tol=1e-10, **kw)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    x0 = s.get_array("Z_iF")
    x, st = s.solve_rhs(rhs, x0)
    xo, conv, it, res = pyoracle.krylov(rp, col, val, rhs, x0, solver, prec, 1e-10, max_iter=20000, restart=m)
    assert conv == 1 and st.converged and st.residue < 1e-10
    assert abs(st.iterations - it) <= max(2, 0.1 * it), (st.iterations, it)
    assert rel(x, xo) < 1e-7
    s.close()


@pytest.mark.parametrize("omega,case", [(1.0, "heat2d"), (1.4, "heat2d"), (1.0, "darcy3d")])
def test_sor_iteration_parity(omega, case):
    """SOR sweeps computed level by level equal the oracle's sequential
    sweeps: the same number of sweeps and the same iterate."""
    print(f"\n[krylov] sor w={omega} {case}", flush=True)
    pr = _SYSTEMS[case]("none")
    s = capi.Solver(pr, solver="sor", preconditioner="none", tol=1e-10, sor_omega=omega)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    x0 = s.get_array("Z_iF")
    x, st = s.solve_rhs(rhs, x0)
    xo, conv, it, res = pyoracle.krylov(rp, col, val, rhs, x0, "sor", "none", 1e-10, max_iter=100000,
                                        restart=int(round(100 * omega)))
    assert conv == 1 and st.converged
    assert abs(st.iterations - it) <= 1, (st.iterations, it)
    assert rel(x, xo) < 1e-9
    s.close()


@pytest.mark.parametrize("solver,prec", [("sor", "none"), ("bicgstab", "ilut"), ("bicgstabl", "ilu-isai"), ("idrs", "amg"),
                                         ("cg", "diagonal"), ("tfqmr", "ilu0"), ("tfqmr", "amg"),
                                         ("cg", "amg"), ("jacobi", "ilu0")])
def test_other_solvers_step_parity(solver, prec):
    """Whole time steps with each solver reach the scheme oracle's solution."""
    print(f"\n[step] {solver} {prec}", flush=True)
    pr = problems.heat_2d(s=12, steps=3)
    s = capi.Solver(pr, solver=solver, preconditioner=prec, tol=1e-13)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        st = s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        assert st.converged
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10, k
    s.close()


def test_ilut_refusals():
    pr = problems.heat_2d(s=6, steps=1)
    with pytest.raises(capi.MHFEMError, match="FP64"):
        capi.Solver(pr, preconditioner="ilut", precision="fp32-refined")
    with pytest.raises(capi.MHFEMError, match="SOR"):
        capi.Solver(pr, solver="sor", preconditioner="ilut")
    s = capi.Solver(pr, preconditioner="ilut")
    with pytest.raises(capi.MHFEMError, match="ILUT"):
        s.step_graph(0.0, pr.dt)
    s.close()


def test_setup_refuses_combinations_that_cannot_run():
    """Checked when the linear system is set up, before any level table or
    preconditioner is built: SOR takes no preconditioner, batches take FP64
    BiCGStab."""
    pr = problems.heat_2d(s=6, steps=1)
    for prec in ("ilu0", "amg", "ilu-isai"):
        with pytest.raises(capi.MHFEMError, match="SOR uses no preconditioner"):
            capi.Solver(pr, solver="sor", preconditioner=prec)
    batch = problems.concatenate([problems.ensemble_member(4, 1, steps=1), problems.ensemble_member(5, 2, steps=1)])
    for solver in ("gmres", "cg"):
        with pytest.raises(capi.MHFEMError, match="batched systems"):
            capi.Solver(batch, solver=solver)
    with pytest.raises(capi.MHFEMError, match="batched systems"):
        capi.Solver(batch, precision="fp32-reliable")


def test_other_solvers_refuse_fp32():
    pr = problems.heat_2d(s=6, steps=1)
    for solver in ("cg", "tfqmr", "jacobi", "bicgstabl", "idrs", "sor"):
        with pytest.raises(capi.MHFEMError, match="BiCGStab"):
            capi.Solver(pr, solver=solver, precision="fp32-refined")


def test_ilu0_step_parity():
    pr = problems.darcy_3d(s=6, steps=2, preconditioner="ilu0")
    s = capi.Solver(pr, tol=1e-13, check_every=1)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        s.step(t, pr.dt)
        it, _ = o.step(t, pr.dt, tol=1e-13)
        assert abs(s.iterations[-1] - it) <= 2
        t += pr.dt
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10


def _max_chunk_imports(rp, col, n, chunk=512):
    """The largest number of values a chunk of `chunk` consecutive faces reads
    from other chunks in either triangle (whole faces: n values each), as the
    library counts them for its single-launch ILU(0) solves."""
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    mx = 0
    for part in (col < rows, col > rows):
        own = rows[part] // n // chunk
        face = col[part] // n
        ext = face // chunk != own
        key = np.unique(own[ext].astype(np.int64) * (len(rp) - 1) + face[ext])
        mx = max(mx, int(np.bincount(key // (len(rp) - 1)).max()) * n)
    return mx


@pytest.mark.parametrize("shuffle", [None, 5])
def test_ilu0_level_launches_when_chunks_import_too_much(shuffle):
    """With the cells in random order the faces are numbered all over the
    mesh, a chunk of 512 faces reads more than the 1,024 values of other
    chunks its LDS holds, and the triangular solves run as one launch per
    dependency level; in the natural order they run as one launch each.
    Both agree with the oracle."""
    pr = problems.darcy_3d(s=8, steps=2, preconditioner="ilu0", shuffle_cells=shuffle)
    o = pyoracle.Oracle(pr)
    rp, col, _, _ = o.csr()
    imports = _max_chunk_imports(np.asarray(rp), np.asarray(col), pr.n)
    assert (imports > 1024) == (shuffle is not None), imports
    print(f"\n[ilu0 chunk imports] shuffle={shuffle} max imports {imports}", flush=True)
    s = capi.Solver(pr, tol=1e-13, check_every=1)
    t = 0.0
    for k in range(pr.steps):
        s.step(t, pr.dt)
        it, _ = o.step(t, pr.dt, tol=1e-13)
        assert abs(s.iterations[-1] - it) <= 2
        t += pr.dt
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10
    s.close()


@pytest.mark.parametrize("n,precision", [(1, "fp64"), (2, "fp64"), (3, "fp64"), (2, "fp32-reliable")])
def test_ilu0_sweeps_on_many_chunks(n, precision):
    """The single-launch ILU(0) triangular solves (one wavefront per chunk of
    64 faces, waits on earlier chunks) on meshes of thousands of chunks, n =
    1, 2, 3, FP64 and FP32: on the first step's system, BiCGStab + ILU(0) to
    1e-10 from the same start takes the oracle's iteration count within +-2
    and reaches its iterate; whole steps solved to 1e-13 stay within 1e-10
    of the oracle."""
    print(f"\n[ilu0 sweeps] n={n} {precision}", flush=True)
    pr = problems.adr_3d(s={1: 20, 2: 16, 3: 12}[n], n=n, steps=2, preconditioner="ilu0")
    if precision == "fp64":
        a = capi.Solver(pr, tol=1e-10, check_every=1)
        a.preIterate(0.0, pr.dt)
        a.assembleLinearSystem(0.0, pr.dt)
        rp, col, val, rhs = a.export_system()
        x0 = a.get_array("Z_iF")
        x, st = a.solve_rhs(rhs, x0)
        xo, conv, it, _ = pyoracle.krylov(rp, col, val, rhs, x0, "bicgstab", "ilu0", 1e-10)
        assert conv == 1 and st.converged
        assert abs(st.iterations - it) <= 2, (st.iterations, it)
        assert rel(x, xo) < 1e-7
        a.close()
    s = capi.Solver(pr, tol=1e-13, precision=precision, inner_tol=0.03)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        st = s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        assert st.converged and st.residue < 1e-13
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10, k
        assert rel(s.get_array("Z_iK"), o.array("Z_iK")) < 1e-10, k
    s.close()


@pytest.mark.parametrize("precision,inner", [("fp32-refined", 1e-4), ("fp32-reliable", 1e-2)])
@pytest.mark.parametrize("prec", ["diagonal", "ilu0"])
def test_fp32_refinement(prec, precision, inner):
    """FP32 inner BiCGStab in FP64 refinement: the residual and the correction
    are FP64, so solved as tightly as the oracle (1e-13) the mixed-precision
    solution meets the same 1e-10 bar against the all-FP64 oracle as the
    FP64 solvers do."""
    pr = problems.adr_3d(s=5, steps=3, preconditioner=prec)
    s = capi.Solver(pr, tol=1e-13, precision=precision, inner_tol=inner)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        st = s.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        assert st.converged and st.residue < 1e-13
        assert st.outer_iterations >= 1
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10, k
        assert rel(s.get_array("Z_iK"), o.array("Z_iK")) < 1e-10, k


@pytest.mark.parametrize("precision,inner", [("fp32-refined", 1e-4), ("fp32-reliable", 1e-2), ("fp64", 0.0)])
def test_face_groups(precision, inner):
    """The iteration in the grouped face order (a permuted copy of the matrix;
    FP32: translated residuals and corrections, FP64: the right-hand side and
    the iterate translated at the start and end of the solve) reaches the
    tolerance and the oracle's solution, with iteration counts close to the
    natural order's."""
    pr = problems.adr_3d(s=5, steps=3, preconditioner="diagonal")
    s = capi.Solver(pr, tol=1e-13, precision=precision, inner_tol=inner, face_groups=1)
    nat = capi.Solver(pr, tol=1e-13, precision=precision, inner_tol=inner)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for k in range(pr.steps):
        st = s.step(t, pr.dt)
        sn = nat.step(t, pr.dt)
        o.step(t, pr.dt, tol=1e-13)
        t += pr.dt
        assert st.converged and st.residue < 1e-13
        assert rel(s.get_array("Z_iF"), o.array("Z_iF")) < 1e-10
        assert abs(st.iterations - sn.iterations) <= max(3, 0.2 * sn.iterations), (st.iterations, sn.iterations)


def test_fp32_reliable_iterations():
    """Reliable updates keep the FP32 recurrence's Krylov space across the
    FP64 residual replacements, so the iteration count stays close to the
    all-FP64 solve's, unlike restarted refinement."""
    pr = problems.adr_3d(s=12, steps=3, preconditioner="diagonal")
    s64 = capi.Solver(pr, tol=1e-10)
    s32 = capi.Solver(pr, tol=1e-10, precision="fp32-reliable", inner_tol=1e-2)
    t = 0.0
    for k in range(pr.steps):
        a = s64.step(t, pr.dt)
        b = s32.step(t, pr.dt)
        t += pr.dt
        print(k, a.iterations, b.iterations, b.outer_iterations)
        assert b.converged and b.residue < 1e-10
        assert b.iterations <= 1.3 * a.iterations + 8
        assert rel(s32.get_array("Z_iF"), s64.get_array("Z_iF")) < 1e-8


def test_nonconvergence_raises():
    pr = problems.darcy_3d(s=5, steps=1, preconditioner="diagonal")
    s = capi.Solver(pr, tol=1e-30, max_iterations=3)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    with pytest.raises(capi.MHFEMError, match="did not converge"):
        s.solveLinearSystem()


def test_unsupported_cells_rejected():
    """Unknown cell shapes are refused before any device work."""
    pts = np.zeros((8, 3))
    cells = np.arange(8, dtype=np.int32)[None, :]
    mesh = capi.MeshDesc(3, 8, 1, pts.ctypes.data, cells.ctypes.data)
    model = capi.ModelDesc(1, 0, 0, 1)
    L = capi.lib()
    for vpc in (5, 6, 3):
        assert not L.mhfem_create_cells(mesh, vpc, model)
        assert b"are not supported" in L.mhfem_last_error(None)


ALL_SOLVERS = ["bicgstab", "gmres", "cg", "tfqmr", "jacobi", "bicgstabl", "idrs", "sor"]


@pytest.mark.parametrize("solver", ALL_SOLVERS)
def test_solvers_on_zero_rhs_and_from_the_solution(solver):
    """Edge cases of every linear solver: a zero right-hand side from a zero
    start needs no iteration and leaves x at zero; started from a solution
    already within the tolerance, a solver stops at once."""
    print(f"\n[krylov] {solver} edge cases", flush=True)
    pr = problems.heat_2d(s=8, steps=1)
    prec = "none" if solver == "sor" else "diagonal"
    s = capi.Solver(pr, solver=solver, preconditioner=prec, tol=1e-10)
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    rp, col, val, rhs = s.export_system()
    n = rhs.size
    x, st = s.solve_rhs(np.zeros(n), np.zeros(n))
    assert st.converged and st.iterations == 0 and not x.any()
    ref = capi.Solver(pr, preconditioner="ilu0", tol=1e-14)
    ref.preIterate(0.0, pr.dt)
    ref.assembleLinearSystem(0.0, pr.dt)
    xs, st_ref = ref.solve_rhs(rhs, np.zeros(n))
    ref.close()
    x, st = s.solve_rhs(rhs, xs)
    assert st.converged and st.iterations == 0, st.iterations
    assert rel(x, xs) < 1e-14
    s.close()
