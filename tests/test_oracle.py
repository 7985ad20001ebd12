"""Pins of the CPU oracle against facts that do not come from its own code.

The reference ships no tests, fixtures or golden vectors (its example meshes
are Git-LFS stubs), so every check here comes from mathematics or from an
independent construction:
  * topology: Euler's relation for the face counts, incidence invariants and
    agreement of two independent implementations of the numbering rule;
  * local mass matrices: exact integration of the RTN0 basis (barycentric
    moments), vertex quadrature for the lumped variant;
  * scheme + assembly + solve + recovery: the full (non-hybridised) mixed
    system assembled densely from the quadrature mass matrices and solved with
    numpy, for n = 1 and n = 2, all boundary-condition types;
  * the linear patch test and mass conservation;
  * ILU(0): (LU)_ij == A_ij on the pattern; Krylov: scipy's direct solve.
"""
import dataclasses
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mhfem_amd import meshgen, problems
import pyoracle


# ---------------------------------------------------------------- topology
@pytest.mark.parametrize("s", [1, 2, 5, 17])
def test_triangle_face_counts(s):
    pts, cells = meshgen.unit_square_triangles(s)
    cf, fc = pyoracle.topology(pts, cells)
    C = cells.shape[0]
    nb = int((fc[:, 1] < 0).sum())
    assert C == 2 * s * s
    assert nb == 4 * s                       # boundary edges of the square
    assert 3 * C == 2 * (fc.shape[0] - nb) + nb  # each cell has 3 faces
    assert fc.shape[0] == 3 * s * s + 2 * s


@pytest.mark.parametrize("s", [1, 2, 3, 6])
def test_tetrahedron_face_counts(s):
    pts, cells = meshgen.unit_cube_kuhn(s)
    cf, fc = pyoracle.topology(pts, cells)
    C = cells.shape[0]
    nb = int((fc[:, 1] < 0).sum())
    assert C == 6 * s ** 3
    assert nb == 12 * s * s                  # 2 triangles per boundary square
    assert 4 * C == 2 * (fc.shape[0] - nb) + nb
    assert fc.shape[0] == 12 * s ** 3 + 6 * s * s


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_topology_invariants(mesh):
    pts, cells = meshgen.unit_square_triangles(7) if mesh == "tri" else meshgen.unit_cube_kuhn(4)
    cf, fc = pyoracle.topology(pts, cells)
    nc, fpc = cells.shape
    # local face e does not contain local vertex e and contains all the others
    fv = {}
    for K in range(nc):
        for e in range(fpc):
            verts = frozenset(int(v) for q, v in enumerate(cells[K]) if q != e)
            E = cf[K, e]
            if E in fv:
                assert fv[E] == verts
            fv[E] = verts
    # every face id appears; first appearances are in increasing order
    flat = cf.ravel()
    _, first = np.unique(flat, return_index=True)
    assert np.all(np.diff(first) > 0)
    # face -> cells and cell -> faces are inverse relations, cells ascending
    for E in range(fc.shape[0]):
        K0, K1 = fc[E]
        assert E in cf[K0]
        if K1 >= 0:
            assert K0 < K1 and E in cf[K1]
    # boundary faces lie on the boundary of the unit domain
    cen = meshgen.face_centroids(pts, cells, cf, fc.shape[0])
    on_bnd = np.any((np.abs(cen) < 1e-12) | (np.abs(cen - 1) < 1e-12), axis=1)
    assert np.array_equal(on_bnd, fc[:, 1] < 0)
    # the numpy implementation of the rule agrees bit for bit
    cf2, fc2 = meshgen.face_numbering(cells, pts.shape[0])
    assert np.array_equal(cf, cf2) and np.array_equal(fc, fc2)


# ---------------------------------------------------------------- mass matrix
def exact_mass(P):
    """M_ef = int_K w_e . w_f, w_e = (x - v_e) / (d |K|), by barycentric moments
    int_K l_k l_l = |K| (1 + delta_kl) / ((d+1)(d+2))."""
    d = P.shape[1]
    n = d + 1
    T = np.ones((n, n)); T[:, 1:] = P
    vol = abs(np.linalg.det(T)) / math.factorial(d)
    mom = vol * (np.ones((n, n)) + np.eye(n)) / ((d + 1) * (d + 2))
    M = np.zeros((n, n))
    for e in range(n):
        for f in range(n):
            De = P - P[e]
            Df = P - P[f]
            M[e, f] = np.einsum("kl,kl->", mom, De @ Df.T) / (d * vol) ** 2
    return M, vol


def lumped_mass(P):
    """vertex quadrature: |K|/(d+1) sum_v w_e(v) . w_f(v)"""
    d = P.shape[1]
    n = d + 1
    T = np.ones((n, n)); T[:, 1:] = P
    vol = abs(np.linalg.det(T)) / math.factorial(d)
    M = np.zeros((n, n))
    for e in range(n):
        for f in range(n):
            M[e, f] = vol / n * sum(np.dot(P[v] - P[e], P[v] - P[f]) for v in range(n)) / (d * vol) ** 2
    return M


def unpack(packed, n):
    M = np.zeros((n, n))
    s = 0
    for c in range(n):
        for r in range(c + 1):
            M[r, c] = M[c, r] = packed[s]
            s += 1
    return M, packed[s:s + n]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("lumping", [0, 1])
def test_local_mass_inverse_matches_quadrature(dim, lumping):
    rng = np.random.default_rng(10 + dim + lumping)
    for trial in range(25):
        P = rng.standard_normal((dim + 1, dim))
        if trial == 0:  # a sliver
            P[-1] = P[0] + 1e-3 * (P[1] - P[0]) + 1e-3 * rng.standard_normal(dim)
        D = rng.uniform(0.1, 10.0)
        out = pyoracle.local_mass_inverse(P, lumping, D)
        Minv, rowsum = unpack(out, dim + 1)
        M = lumped_mass(P) if lumping else exact_mass(P)[0]
        ref = D * np.linalg.inv(M)
        # inversion error grows with the condition number (slivers)
        tol = max(1e-12, 1e-14 * np.linalg.cond(M))
        assert np.abs(Minv - ref).max() / np.abs(ref).max() < tol
        assert np.allclose(rowsum, Minv.sum(axis=1), rtol=1e-12, atol=1e-12 * np.abs(Minv).max())


def test_mass_matrix_spd_and_unit_triangle():
    # unit right triangle: the integrals of the RTN0 basis by hand give
    # M = [[1/2, -1/6, 1/6]... ] only through the moments; check SPD instead
    P = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    out = pyoracle.local_mass_inverse(P, 0, 1.0)
    Minv, _ = unpack(out, 3)
    assert np.all(np.linalg.eigvalsh(Minv) > 0)
    M, vol = exact_mass(P)
    # w_0 = (x - v0)/(2|K|) integrates |x|^2/(4|K|^2): int_K x^2+y^2 = 1/6 -> M00 = (1/6)/(4/4) = 1/6
    assert abs(M[0, 0] - 1.0 / 6.0) < 1e-14
    assert np.allclose(np.linalg.inv(Minv), M, rtol=1e-12)


# ---------------------------------------------------------------- box cells
def box_mass(P, lumping):
    """M_ef = int_K w_e . w_f for the RTN0 basis of an axis-aligned box (an
    edge or a rectangle, vertices in VTK order): w_e = +-(x_a - c) / |K| e_a,
    with a the axis of face e's normal and c the coordinate of the opposite
    face, so that w_e . n_e = 1 / |E_e| on face e and 0 on the others.  Exact
    2-point Gauss quadrature per axis, or vertex (trapezoidal) quadrature when
    lumped.  Returns (M, |K|)."""
    d = P.shape[1]
    lo, hi = P[0], P[{1: 1, 2: 2, 3: 6}[d]]
    h = hi - lo
    vol = float(np.prod(np.abs(h)))
    # local faces: (axis, the face's side: +1 at hi, -1 at lo)
    faces = {1: [(0, -1), (0, +1)], 2: [(1, -1), (0, +1), (1, +1), (0, -1)],
             3: [(2, -1), (1, -1), (0, +1), (1, +1), (0, -1), (2, +1)]}[d]

    def w(e, x):
        a, side = faces[e]
        c = lo[a] if side > 0 else hi[a]
        out = np.zeros(d)
        out[a] = (x[a] - c) / vol  # points out of the face at either end
        return out

    if lumping:
        corners = [lo + h * np.array(bits) for bits in itertools.product([0, 1], repeat=d)]
        pts, wts = corners, [vol / len(corners)] * len(corners)
    else:
        g = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
        pts = [lo + h * np.array(t) for t in itertools.product(g, repeat=d)]
        wts = [vol / len(pts)] * len(pts)
    n = len(faces)
    M = np.zeros((n, n))
    for x, wt in zip(pts, wts):
        for e in range(n):
            for f in range(n):
                M[e, f] += wt * w(e, x) @ w(f, x)
    return M, vol


def box_b_from_storage(dim, lumping, st):
    """The full b_ijKef of one (i,j,K) from its box storage (MassMatrix.h:57-71,
    118-132, 190-205, 250-276)."""
    fpc = 2 * dim
    d = {1: [0, 0], 2: [0, 1, 0, 1], 3: [2, 1, 0, 1, 0, 2]}[dim]
    B = np.zeros((fpc, fpc))
    for e in range(fpc):
        for f in range(fpc):
            if lumping:
                B[e, f] = st[d[e]] if e == f else 0.0
            elif d[e] == d[f]:
                B[e, f] = 2 * st[d[e]] if e == f else st[d[e]]
    return B


def _box_mesh(dim, s, rng):
    """A box mesh with uneven, anisotropic spacing (hx != hy in every cell)."""
    if dim == 1:
        pts, cells = meshgen.unit_interval(s)
        xs = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, s - 1)), [1.0]])
        return xs[:, None], cells
    pts, cells = meshgen.unit_square_quads(s) if dim == 2 else meshgen.unit_cube_hexes(s)
    lines = [np.concatenate([[0.0], np.sort(rng.uniform(0, L, s - 1)), [L]]) for L in (1.0, 0.7, 1.3)[:dim]]
    idx = np.rint(pts * s).astype(int)
    pts = np.stack([lines[a][idx[:, a]] for a in range(dim)], axis=1)
    return pts, cells


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("lumping", [0, 1])
def test_box_mass_matrix_matches_quadrature(dim, lumping):
    rng = np.random.default_rng(40 + dim + 2 * lumping)
    pr = {1: lambda: problems.adr_1d(s=4, n=2, lumping=lumping),
          2: lambda: problems.adr_quads(s=3, n=2, lumping=lumping),
          3: lambda: problems.adr_hexes(s=2, n=2, lumping=lumping)}[dim]()
    pts, cells = _box_mesh(dim, {1: 4, 2: 3, 3: 2}[dim], rng)
    cf, fc = meshgen.face_numbering(cells, pts.shape[0], dim)
    C = cells.shape[0]
    D = rng.uniform(0.1, 10.0, (2, 2, C))
    pr = dataclasses.replace(pr, points=pts, cells=cells, cell_faces=cf, face_cells=fc,
                             coeffs=dict(pr.coeffs, D_ijK=D.ravel()))
    o = pyoracle.Oracle(pr)
    o.pre_iterate(0.0, pr.dt)
    st = o.array("b_ijK_storage").reshape(2, 2, dim, C)
    meas = o.array("cell_measure")
    for K in range(C):
        M, vol = box_mass(pts[cells[K]], lumping)
        assert abs(meas[K] - vol) < 1e-14 * vol
        for i in range(2):
            for j in range(2):
                ref = D[i, j, K] * np.linalg.inv(M)
                got = box_b_from_storage(dim, lumping, st[i, j, :, K])
                assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), (K, i, j)


# ---------------------------------------------------------------- brute force
def _cell_mass(pr, K):
    """(M, |K|) of cell K by quadrature of the basis."""
    P = pr.points[pr.cells[K]]
    if pr.box:
        return box_mass(P, pr.lumping)
    if not pr.lumping:
        return exact_mass(P)
    T = np.ones((pr.fpc, pr.fpc)); T[:, 1:] = P
    return lumped_mass(P), abs(np.linalg.det(T)) / math.factorial(P.shape[1])


def _face_measure(pr, K, e):
    verts = pr.points[pr.cells[K][meshgen.face_vertex_table(pr.dim, pr.cells.shape[1])[e]]]
    if pr.dim == 1:
        return 1.0
    if pr.dim == 2:
        return np.linalg.norm(verts[1] - verts[0])
    if len(verts) == 4:  # a hexahedron's rectangle
        return np.linalg.norm(verts[1] - verts[0]) * np.linalg.norm(verts[3] - verts[0])
    return 0.5 * np.linalg.norm(np.cross(verts[1] - verts[0], verts[2] - verts[0]))


def _mixed_reference(pr, o_pts, tau, Z_old, Z_F_old, m_upw, bc_t):
    """Dense full mixed system (velocities, cell and face unknowns) for one
    time step, explicit upwind advection, from quadrature mass matrices."""
    pts, cells = pr.points, pr.cells
    cf, fc = pr.cell_faces, pr.face_cells
    n, C, F, fpc = pr.n, pr.num_cells, pr.num_faces, pr.fpc
    dim = pts.shape[1]
    co = {k: pr.coefficient_array(k) for k in problems.COEFFICIENTS}
    Nij = co["N_ijK"].reshape(n, n, C)
    Dij = co["D_ijK"].reshape(n, n, C)
    rij = co["r_ijK"].reshape(n, n, C)
    fi = co["f_iK"].reshape(n, C)
    wi = co["w_iKe"].reshape(n, fpc, C)
    vel = (co["a_ijKe"] + co["u_ijKe"]).reshape(n, n, fpc, C)
    uu = co["u_ijKe"].reshape(n, n, fpc, C)
    nv = n * C * fpc
    iv = lambda i, K, e: (i * C + K) * fpc + e
    ik = lambda i, K: nv + i * C + K
    iF = lambda i, E: nv + n * C + i * F + E
    N = nv + n * C + n * F
    A = np.zeros((N, N))
    b = np.zeros(N)
    Minv = []
    vols = []
    for K in range(C):
        M, vol = _cell_mass(pr, K)
        Minv.append(np.linalg.inv(M))
        vols.append(vol)
    # upwinded Z on faces (Solver.hpp:743-759): from the old state
    def zupw(i, j, E):
        K1, K2 = fc[E]
        e1 = list(cf[K1]).index(E)
        apu = vel[i, j, e1, K1]
        if apu > 0:
            return Z_old[j, K1]
        if apu == 0:
            return 0.0
        return Z_old[j, K2] if K2 >= 0 else Z_F_old[j, E]
    r = 0
    for K in range(C):
        for i in range(n):
            for e in range(fpc):
                # v_iKe - sum_j D_ij Minv (Z_jK - Z_jF) = w_iKe
                A[r, iv(i, K, e)] = 1.0
                for j in range(n):
                    for f in range(fpc):
                        bb = Dij[i, j, K] * Minv[K][e, f]
                        A[r, ik(j, K)] -= bb
                        A[r, iF(j, cf[K, f])] += bb
                b[r] = wi[i, e, K]
                r += 1
    for K in range(C):
        for i in range(n):
            # sum_e m v_iKe + sum_j (|K| N_ij / tau + |K| r_ij - sum_e u_ijKe) Z_jK = rhs
            for e in range(fpc):
                A[r, iv(i, K, e)] += m_upw[cf[K, e], i]
            rhs = vols[K] * fi[i, K]
            for j in range(n):
                A[r, ik(j, K)] += vols[K] * Nij[i, j, K] / tau + vols[K] * rij[i, j, K] - uu[i, j, :, K].sum()
                rhs += vols[K] * Nij[i, j, K] / tau * Z_old[j, K]
                for e in range(fpc):
                    rhs -= vel[i, j, e, K] * zupw(i, j, cf[K, e])
            b[r] = rhs
            r += 1
    for E in range(F):
        K0, K1 = fc[E]
        e0 = list(cf[K0]).index(E)
        meas = _face_measure(pr, K0, e0)
        for i in range(n):
            if K1 >= 0:
                e1 = list(cf[K1]).index(E)
                A[r, iv(i, K0, e0)] = 1.0
                A[r, iv(i, K1, e1)] = 1.0
            else:
                tag = pr.tags[i * F + E]
                if tag == problems.FIXED_VALUE:
                    A[r, iF(i, E)] = 1.0
                    b[r] = bc_t[i * F + E]
                elif tag == problems.FIXED_FLUX_NO_ADVECTION:
                    A[r, iv(i, K0, e0)] = 1.0
                    b[r] = pr.bc_values[i * F + E] * meas
                elif tag == problems.FIXED_FLUX:
                    A[r, iv(i, K0, e0)] = m_upw[E, i]
                    b[r] = pr.bc_values[i * F + E] * meas - sum(vel[i, j, e0, K0] * zupw(i, j, E) for j in range(n))
                elif tag == problems.ADVECTIVE_OUTFLOW:
                    if vel[i, i, e0, K0] >= 0:
                        A[r, iv(i, K0, e0)] = 1.0
                    else:
                        A[r, iF(i, E)] = 1.0
                        b[r] = Z_old[i, K0]
            r += 1
    assert r == N
    x = np.linalg.solve(A, b)
    # the fluxes v_iKe in the reference's (i, e, K) layout
    v = x[:nv].reshape(n, C, fpc).transpose(0, 2, 1)
    return x[nv + n * C:].reshape(n, F), x[nv:nv + n * C].reshape(n, C), v


def _tiny_problem(kind, seed, mobility=False):
    """A few-cell problem of every coefficient kind with random data and all
    four boundary types.  ``mobility``: the upwinded mobility model
    (do_mobility_upwind) with a random, non-uniform m_iK and boundary
    mobility, otherwise the constant-mobility model (m_iE_upw = 1)."""
    rng = np.random.default_rng(seed)
    if kind == "tri_n1":
        pr = problems.heat_2d(s=3, steps=2)
        pts = pr.points.copy()
        inner = np.all((pts > 1e-12) & (pts < 1 - 1e-12), axis=1)
        pts[inner] += 0.05 * rng.standard_normal((inner.sum(), 2))  # unstructured interior
        pr = dataclasses.replace(pr, points=pts)
    elif kind == "tet_n1":
        pr = problems.darcy_3d(s=2, steps=2, preconditioner="diagonal")
    elif kind == "adr_n3":
        pr = problems.adr_3d(s=2, steps=2, preconditioner="diagonal", n=3)
    elif kind == "tri_adr_n2":
        pr = problems.adr_2d(s=3, n=2, steps=2)
        pts = pr.points.copy()
        inner = np.all((pts > 1e-12) & (pts < 1 - 1e-12), axis=1)
        pts[inner] += 0.05 * rng.standard_normal((inner.sum(), 2))
        pr = dataclasses.replace(pr, points=pts)
    elif kind.startswith(("quad", "edge", "hex")):
        dim = {"edge": 1, "quad": 2, "hex_": 3}[kind[:4]]
        n = int(kind[-1])
        lumping = int("lumped" in kind)
        pr = {1: lambda: problems.adr_1d(s=5, n=n, lumping=lumping),
              2: lambda: problems.adr_quads(s=3, n=n, lumping=lumping),
              3: lambda: problems.adr_hexes(s=2, n=n, lumping=lumping)}[dim]()
        pts, cells = _box_mesh(dim, {1: 5, 2: 3, 3: 2}[dim], rng)
        cf, fc = meshgen.face_numbering(cells, pts.shape[0], dim)
        pr = dataclasses.replace(pr, points=pts, cells=cells, cell_faces=cf, face_cells=fc,
                                 steps=2)
        pr.coeffs["a_ijKe"] = 0.3 * rng.standard_normal(n * n * pr.fpc * pr.num_cells)
    else:
        pr = problems.adr_3d(s=2, steps=2, preconditioner="diagonal")
    n, C, F, fpc = pr.n, pr.num_cells, pr.num_faces, pr.fpc
    co = dict(pr.coeffs)
    co["f_iK"] = rng.uniform(-1, 1, n * C)
    co["r_ijK"] = (np.eye(n)[:, :, None] * rng.uniform(0, 0.5, (1, 1, C))).ravel()
    co["N_ijK"] = (np.eye(n)[:, :, None] * rng.uniform(0.5, 2, (1, 1, C))).ravel()
    if n >= 2:
        K = rng.uniform(0.5, 2, C)
        co["D_ijK"] = np.stack([K if i == j else 0.1 * K for i in range(n) for j in range(n)]).ravel()
    else:
        co["D_ijK"] = rng.uniform(0.5, 2.0, C)
    co["w_iKe"] = 0.1 * rng.standard_normal(n * fpc * C)
    tags = np.array(pr.tags, copy=True)
    vals = np.zeros(n * F)
    dirv = np.zeros(n * F)
    bnd = np.nonzero(pr.face_cells[:, 1] < 0)[0]
    for i in range(n):
        for q, E in enumerate(bnd):
            tags[i * F + E] = [problems.FIXED_VALUE, problems.FIXED_FLUX_NO_ADVECTION,
                               problems.FIXED_FLUX, problems.ADVECTIVE_OUTFLOW][(q + i) % 4]
            vals[i * F + E] = rng.uniform(-1, 1)
            dirv[i * F + E] = rng.uniform(-1, 1)
    z0 = rng.uniform(0, 1, n * C)
    if mobility:
        co["m_iK"] = rng.uniform(0.5, 2.0, n * C)
        return dataclasses.replace(pr, coeffs=co, tags=tags, bc_values=vals, bc_dirichlet=dirv, Z_iK0=z0,
                                   do_mobility_upwind=1, bc_mobility=rng.uniform(0.5, 2.0, n * F))
    return dataclasses.replace(pr, coeffs=co, tags=tags, bc_values=vals, bc_dirichlet=dirv, Z_iK0=z0,
                               do_mobility_upwind=0, bc_mobility=np.ones(n * F))


@pytest.mark.parametrize("kind", ["tri_n1", "tet_n1", "adr_n2", "adr_n3", "tri_adr_n2",
                                  "quad_n1", "quad_n2", "quad_lumped_n3", "edge_n1", "edge_n2",
                                  "edge_lumped_n2", "hex_n1", "hex_n2", "hex_lumped_n2"])
def test_brute_force_mixed_system(kind):
    pr = _tiny_problem(kind, 3)
    o = pyoracle.Oracle(pr)
    n, C, F = pr.n, pr.num_cells, pr.num_faces
    t = 0.0
    for step in range(2):
        Z_old = o.array("Z_iK").reshape(n, C).copy()
        ZF_old = o.array("Z_iF").reshape(F, n).T.copy()
        m_upw = o.array("m_iE_upw").reshape(F, n).copy()
        o.step(t, pr.dt, tol=1e-14)
        ZF, ZK, _ = _mixed_reference(pr, None, pr.dt, Z_old, ZF_old, m_upw, pr.bc_dirichlet)
        got_F = o.array("Z_iF").reshape(F, n).T
        got_K = o.array("Z_iK").reshape(n, C)
        assert np.abs(got_F - ZF).max() / np.abs(ZF).max() < 1e-9, step
        assert np.abs(got_K - ZK).max() / np.abs(ZK).max() < 1e-9, step
        t += pr.dt


def _face_average(pr, Z):
    """K11 (Solver.hpp:319-339): a face's initial trace is the mean of its
    cells' values (the one cell on the boundary)."""
    K0, K1 = pr.face_cells[:, 0], pr.face_cells[:, 1]
    return np.where(K1 < 0, Z[:, K0], 0.5 * (Z[:, K0] + Z[:, np.maximum(K1, 0)]))


def _upwind_mobility(pr, v, m, bc_mob):
    """Solver.hpp:663-727 applied to fluxes v[i, e, K]: on an interior face
    the mobility of the cell the flux of species i leaves (vel = v_iK1e1 -
    v_iK2e2 > 0: K1 = cellIndexes[0], < 0: K2, == 0: the mean); on a boundary
    face the boundary mobility when any species flows in, else the cell's.
    Returns (m_iE_upw as [E, i], mask of decisions taken within rounding of a
    tie, which a test cannot hold to one side)."""
    n, C, F = pr.n, pr.num_cells, pr.num_faces
    cf, fc = pr.cell_faces, pr.face_cells
    m = m.reshape(n, C)
    bcm = np.asarray(bc_mob).reshape(n, F)
    eps = 1e-9 * max(np.abs(v).max(), 1e-300)
    out = np.empty((F, n))
    near = np.zeros((F, n), bool)
    for E in range(F):
        K1, K2 = fc[E]
        e1 = list(cf[K1]).index(E)
        for i in range(n):
            if K2 < 0:
                vj = v[:, e1, K1]
                out[E, i] = bcm[i, E] if np.any(vj < 0) else m[i, K1]
                near[E, i] = bool(np.any((np.abs(vj) < eps) & (vj != 0)))
            else:
                e2 = list(cf[K2]).index(E)
                vel = v[i, e1, K1] - v[i, e2, K2]
                out[E, i] = m[i, K1] if vel > 0 else m[i, K2] if vel < 0 else 0.5 * (m[i, K1] + m[i, K2])
                near[E, i] = vel != 0 and abs(vel) < eps
    return out, near


@pytest.mark.parametrize("kind", ["tri_n1", "tet_n1", "adr_n2", "adr_n3", "tri_adr_n2", "quad_n2",
                                  "edge_n2", "hex_n2", "hex_lumped_n2"])
def test_brute_force_fluxes_and_mobility_upwind(kind):
    """K4 and K10 against the dense mixed system, over three steps of a
    problem with a non-uniform mobility m_iK, a boundary mobility different
    from it and random boundary data: the brute-force chain starts from the
    initial cell values and the face averages, takes its mobility upwind from
    its own fluxes of the step before (zero before the first step), and
    solves for the fluxes v_iKe together with the traces.  The oracle's
    m_iE_upw must equal the brute-force choice, and its velocities and
    traces the brute-force ones."""
    pr = _tiny_problem(kind, 11, mobility=True)
    # boundary types whose fluxes are away from zero for every species, so
    # that the any-species inflow test is decided by the data, not by rounding
    # (AdvectiveOutflow rows carry a zero diffusive flux)
    tags = np.array(pr.tags, copy=True)
    vals = np.array(pr.bc_values, copy=True)
    rng = np.random.default_rng(12)
    ao = tags == problems.ADVECTIVE_OUTFLOW
    tags[ao] = problems.FIXED_FLUX
    flux = tags != problems.FIXED_VALUE
    vals[flux] = rng.choice([-1.0, 1.0], flux.sum()) * rng.uniform(0.2, 1.0, flux.sum())
    pr = dataclasses.replace(pr, tags=tags, bc_values=vals)
    o = pyoracle.Oracle(pr)
    n, C, F, fpc = pr.n, pr.num_cells, pr.num_faces, pr.fpc
    m = pr.coefficient_array("m_iK").reshape(n, C)
    Z_old = np.asarray(pr.Z_iK0).reshape(n, C)
    ZF_old = _face_average(pr, Z_old)
    assert np.abs(o.array("Z_iF").reshape(F, n).T - ZF_old).max() < 1e-15
    v_prev = np.zeros((n, fpc, C))
    bnd = pr.face_cells[:, 1] < 0
    seen = {"K1": 0, "K2": 0, "tie": 0, "inflow_by_other": 0, "inflow": 0}
    t = 0.0
    for step in range(3):
        m_upw, near = _upwind_mobility(pr, v_prev, m, pr.bc_mobility)
        assert not near.any(), step
        o.pre_iterate(t, pr.dt)
        got = o.array("m_iE_upw").reshape(F, n)
        assert np.array_equal(got, m_upw), (step, np.argwhere(got != m_upw)[:5])
        for E in range(F):
            K1, K2 = pr.face_cells[E]
            for i in range(n):
                if K2 >= 0:
                    seen["K1" if m_upw[E, i] == m[i, K1] != m[i, K2] else "K2" if m_upw[E, i] == m[i, K2] != m[i, K1]
                         else "tie"] += 1
                else:
                    e1 = list(pr.cell_faces[K1]).index(E)
                    if np.any(v_prev[:, e1, K1] < 0):
                        seen["inflow"] += 1
                        seen["inflow_by_other"] += int(v_prev[i, e1, K1] >= 0)
        o.assemble(t, pr.dt)
        o.solve(tol=1e-14)
        o.post_iterate(t, pr.dt)
        ZF, ZK, v = _mixed_reference(pr, None, pr.dt, Z_old, ZF_old, m_upw, pr.bc_dirichlet)
        got_F = o.array("Z_iF").reshape(F, n).T
        got_K = o.array("Z_iK").reshape(n, C)
        got_v = o.array("v_iKe").reshape(n, fpc, C)
        assert np.abs(got_F - ZF).max() / np.abs(ZF).max() < 1e-9, step
        assert np.abs(got_K - ZK).max() / np.abs(ZK).max() < 1e-9, step
        assert np.abs(got_v - v).max() / np.abs(v).max() < 1e-9, step
        v_prev, Z_old, ZF_old = v, ZK, ZF
        t += pr.dt
    # every branch of the rule was taken: both upwind cells, the tie of the
    # first step, boundary inflow, and (with several species) inflow of
    # species i decided by another species' flux
    assert seen["K1"] > 0 and seen["K2"] > 0 and seen["tie"] > 0 and seen["inflow"] > 0, seen
    if n > 1 and pr.dim > 1:
        assert seen["inflow_by_other"] > 0, seen


def _area_normals(pr):
    """|E_e| n_{K,e} of every local face of a simplex, from the vertices:
    face e is opposite vertex e, and the normal points away from it."""
    P = pr.points[pr.cells]  # (C, d+1, d)
    C, nv, d = P.shape
    out = np.zeros((C, nv, d))
    for e in range(nv):
        others = [q for q in range(nv) if q != e]
        a = P[:, others[0]]
        if d == 2:
            t = P[:, others[1]] - a
            nrm = np.stack([t[:, 1], -t[:, 0]], axis=1)
        else:
            nrm = 0.5 * np.cross(P[:, others[1]] - a, P[:, others[2]] - a)
        flip = np.einsum("kd,kd->k", nrm, P[:, e] - a) > 0
        nrm[flip] *= -1
        out[:, e] = nrm
    return out


@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_velocities_and_upwind_of_a_linear_field(mesh):
    """K10 and K4 on the linear patch test.  Steady diffusion (N = 0) with
    Dirichlet data from Z = g.x on the whole boundary is reproduced exactly by
    RTN0, and so is its constant velocity: the oracle's v_iKe must equal
    -D g . n_{K,e} |E_e| with the normals computed here from the vertices.
    Then, with a non-uniform mobility, the next step's m_iE_upw must take on
    every interior face the mobility of the cell the flow -D g leaves, and on
    the boundary the boundary mobility exactly where the flow enters."""
    pr = problems.heat_2d(s=5) if mesh == "tri" else problems.darcy_3d(s=3, preconditioner="diagonal")
    F, C = pr.num_faces, pr.num_cells
    g = np.array([0.7, -0.45, 0.3])[:pr.dim]
    D = 2.5
    fcen = meshgen.face_centroids(pr.points, pr.cells, pr.cell_faces, F)
    co = dict(pr.coeffs, N_ijK=np.array(0.0), D_ijK=np.array(D), m_iK=np.array(1.0))
    rng = np.random.default_rng(4)
    pr = dataclasses.replace(pr, coeffs=co, tags=np.full(F, problems.FIXED_VALUE, np.uint8),
                             bc_values=np.zeros(F), bc_dirichlet=0.3 + fcen @ g, Z_iK0=np.zeros(C),
                             do_mobility_upwind=1, bc_mobility=rng.uniform(0.5, 2.0, F))
    o = pyoracle.Oracle(pr)
    o.step(0.0, 1.0, tol=1e-14)
    nrm = _area_normals(pr)  # (C, fpc, d)
    expect = -D * np.einsum("ked,d->ek", nrm, g)
    v = o.array("v_iKe").reshape(pr.fpc, C)
    assert np.abs(v - expect).max() < 1e-11 * np.abs(expect).max()
    # the mobility upwind of the next step follows the flow
    m = rng.uniform(0.5, 2.0, C)
    o.array("m_iK")[:] = m
    o.pre_iterate(1.0, 1.0)
    got = o.array("m_iE_upw")
    flow = -D * g
    for E in range(F):
        K1, K2 = pr.face_cells[E]
        e1 = list(pr.cell_faces[K1]).index(E)
        out_of_K1 = flow @ nrm[K1, e1]
        assert abs(out_of_K1) > 1e-3 * np.linalg.norm(nrm[K1, e1])  # no face lies along the flow
        if K2 >= 0:
            assert got[E] == (m[K1] if out_of_K1 > 0 else m[K2]), E
        else:
            assert got[E] == (pr.bc_mobility[E] if out_of_K1 < 0 else m[K1]), E


def test_mobility_upwind_of_assigned_fluxes():
    """K4 on fluxes set directly: species 0 and 1 carry the fluxes of two
    constant fields of different directions (v = u_i . n |E|, normals from
    the vertices), species 2 carries none.  Interior faces take the mobility
    of the cell each species' field leaves, the still species the mean of the
    two cells; a boundary face takes the boundary mobility for every species
    as soon as one species' field enters there (Solver.hpp:680-690)."""
    pr = _tiny_problem("adr_n3", 2, mobility=True)
    n, C, F = pr.n, pr.num_cells, pr.num_faces
    o = pyoracle.Oracle(pr)
    nrm = _area_normals(pr)
    u = np.array([[1.0, 0.37, 0.21], [-0.3, 0.8, -0.45], [0.0, 0.0, 0.0]])
    v = np.einsum("id,ked->iek", u, nrm)
    o.array("v_iKe")[:] = v.ravel()
    o.pre_iterate(0.0, pr.dt)
    got = o.array("m_iE_upw").reshape(F, n)
    m = pr.coefficient_array("m_iK").reshape(n, C)
    bcm = pr.bc_mobility.reshape(n, F)
    other_decides = 0
    for E in range(F):
        K1, K2 = pr.face_cells[E]
        e1 = list(pr.cell_faces[K1]).index(E)
        out = u @ nrm[K1, e1]
        for i in range(n):
            if K2 >= 0:
                ref = 0.5 * (m[i, K1] + m[i, K2]) if i == 2 else (m[i, K1] if out[i] > 0 else m[i, K2])
            else:
                inflow = np.any(out < 0)
                ref = bcm[i, E] if inflow else m[i, K1]
                other_decides += int(inflow and out[i] >= 0)
            assert got[E, i] == ref, (E, i)
    assert other_decides > 0


# ---------------------------------------------------------------- invariants
@pytest.mark.parametrize("mesh", ["tri", "tet"])
def test_linear_patch_test(mesh):
    """Steady diffusion (N=0) with Dirichlet data from a linear function: the
    face traces equal the function at face centroids and the cell values at
    cell centroids (RTN0 MHFEM is exact on linear solutions)."""
    pr = problems.heat_2d(s=6) if mesh == "tri" else problems.darcy_3d(s=3, preconditioner="diagonal")
    F = pr.num_faces
    g = lambda x: 0.3 + x @ np.arange(1, x.shape[1] + 1) * 0.7
    fc = meshgen.face_centroids(pr.points, pr.cells, pr.cell_faces, F)
    tags = np.full(F, problems.FIXED_VALUE, np.uint8)
    tags[pr.face_cells[:, 1] >= 0] = problems.FIXED_VALUE
    co = dict(pr.coeffs)
    co["N_ijK"] = np.array(0.0)
    co["D_ijK"] = np.array(2.5)
    pr = dataclasses.replace(pr, coeffs=co, tags=tags, bc_values=np.zeros(F), bc_dirichlet=g(fc),
                             Z_iK0=np.zeros(pr.num_cells), do_mobility_upwind=0)
    o = pyoracle.Oracle(pr)
    o.step(0.0, 1.0, tol=1e-14)
    assert np.abs(o.array("Z_iF") - g(fc)).max() < 1e-10
    cc = meshgen.cell_centroids(pr.points, pr.cells)
    assert np.abs(o.array("Z_iK") - g(cc)).max() < 1e-10


def test_conservation_zero_flux():
    pr = problems.darcy_3d(s=3, preconditioner="diagonal", steps=4)
    F = pr.num_faces
    rng = np.random.default_rng(0)
    pr = dataclasses.replace(pr, tags=np.full(F, problems.FIXED_FLUX_NO_ADVECTION, np.uint8),
                             Z_iK0=rng.uniform(0, 1, pr.num_cells))
    o = pyoracle.Oracle(pr)
    vol = o.array("cell_measure").copy()
    mass0 = (vol * o.array("Z_iK")).sum()
    t = 0.0
    for _ in range(pr.steps):
        o.step(t, pr.dt, tol=1e-14)
        t += pr.dt
        assert abs((vol * o.array("Z_iK")).sum() - mass0) < 1e-10 * abs(mass0)


def test_csr_pattern_invariants():
    for pr in (problems.adr_3d(s=2), problems.heat_2d(s=4)):
        o = pyoracle.Oracle(pr)
        rp, col, val, rhs = o.csr()
        n, F, fpc = pr.n, pr.num_faces, pr.fpc
        cf, fc = pr.cell_faces, pr.face_cells
        for E in range(F):
            for i in range(n):
                r = E * n + i
                cols = col[rp[r]:rp[r + 1]]
                assert np.all(np.diff(cols) > 0)
                K0, K1 = fc[E]
                if K1 >= 0:
                    faces = set(cf[K0]) | set(cf[K1])
                    assert len(cols) == (2 * fpc - 1) * n
                elif pr.tags[i * F + E] == problems.FIXED_VALUE:
                    assert list(cols) == [r]
                    continue
                else:
                    faces = set(cf[K0])
                    assert len(cols) == fpc * n
                assert set(cols) == {Fc * n + j for Fc in faces for j in range(n)}


# ---------------------------------------------------------------- linear algebra
def _mhfem_system(prec="diagonal"):
    pr = problems.darcy_3d(s=4, preconditioner=prec)
    o = pyoracle.Oracle(pr)
    o.pre_iterate(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    rp, col, val, rhs = o.csr()
    return np.array(rp), np.array(col), np.array(val), np.array(rhs), o


def test_ilu0_reproduces_pattern():
    rp, col, val, rhs, _ = _mhfem_system()
    lu = pyoracle.ilu0(rp, col, val)
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    Lm = sp.csr_matrix((np.where(col < rows, lu, 0.0), col, rp), shape=(n, n)) + sp.identity(n)
    Um = sp.csr_matrix((np.where(col >= rows, lu, 0.0), col, rp), shape=(n, n))
    P = (Lm @ Um).tocsr()
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    on = np.asarray(P[rows, col]).ravel()
    assert np.abs(on - val).max() < 1e-12 * np.abs(val).max()
    # and the apply solves L U y = x
    x = np.random.default_rng(1).standard_normal(n)
    y = pyoracle.ilu0_apply(rp, col, lu, x)
    assert np.linalg.norm(P @ y - x) < 1e-11 * np.linalg.norm(x)
    del A


@pytest.mark.parametrize("solver", ["bicgstab", "gmres", "tfqmr", "jacobi"])
@pytest.mark.parametrize("prec", ["none", "diagonal", "ilu0"])
def test_krylov_against_direct_solve(solver, prec):
    rp, col, val, rhs, o = _mhfem_system(prec)
    n = rp.size - 1
    A = sp.csr_matrix((val, col, rp), shape=(n, n)).tocsc()
    xd = spla.spsolve(A, rhs)
    x, conv, it, res = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), solver, prec, 1e-13)
    assert conv == 1 and res < 1e-13
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


@pytest.mark.parametrize("prec", ["none", "diagonal", "ilu0"])
def test_cg_on_spd_system(prec):
    """CG on the symmetric positive definite 2-D Laplacian (5-point, Dirichlet),
    whose solution scipy gives directly; and CG's residuals are the Lanczos
    ones, so without a preconditioner it takes at most as many iterations as
    the system has distinct eigenvalues (here a small grid, exact arithmetic
    aside)."""
    m = 12
    T = sp.diags([-np.ones(m - 1), 4 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    A = (sp.kron(sp.identity(m), T) + sp.kron(sp.diags([-np.ones(m - 1), -np.ones(m - 1)], [-1, 1]),
                                                 sp.identity(m))).tocsr()
    A.sort_indices()
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    xd = spla.spsolve(A.tocsc(), b)
    x, conv, it, res = pyoracle.krylov(A.indptr, A.indices, A.data, b, np.zeros(A.shape[0]), "cg", prec, 1e-12)
    assert conv == 1 and res < 1e-12
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    if prec == "none":
        assert it <= len(np.unique(np.round(np.linalg.eigvalsh(A.toarray()), 8))) + 1


def test_jacobi_iteration_count_matches_spectral_radius():
    """The stationary iteration x += M^{-1}(b - A x) contracts the error by
    rho(I - M^{-1}A) per step: on the diagonal preconditioner the number of
    iterations to 1e-8 is close to log(1e-8)/log(rho)."""
    rp, col, val, rhs, _ = _mhfem_system("diagonal")
    n = rp.size - 1
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    rho = max(abs(np.linalg.eigvals(np.eye(n) - A.toarray())))
    assert rho < 1
    x, conv, it, res = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), "jacobi", "diagonal", 1e-8,
                                       max_iter=100000)
    assert conv == 1
    expect = np.log(1e-8) / np.log(rho)
    assert 0.3 * expect < it < 1.5 * expect, (it, expect)


def test_rtn0_projection_of_constant_field():
    """a_iiKe = |E| a.n_{K,e}: fluxes of a constant field sum to zero over each
    cell (divergence-free) and are opposite on the two sides of a face."""
    pr = problems.adr_3d(s=3)
    C, F = pr.num_cells, pr.num_faces
    a = np.asarray(pr.coeffs["a_ijKe"]).reshape(2, 2, 4, C)[0, 0]
    assert np.abs(a.sum(axis=0)).max() < 1e-14
    cf, fc = pr.cell_faces, pr.face_cells
    inner = np.nonzero(fc[:, 1] >= 0)[0]
    for E in inner[:200]:
        K0, K1 = fc[E]
        e0 = list(cf[K0]).index(E)
        e1 = list(cf[K1]).index(E)
        assert abs(a[e0, K0] + a[e1, K1]) < 1e-14


@pytest.mark.parametrize("prec", ["none", "ilu0"])
def test_bicgstabl_with_l1_is_bicgstab_and_idr1_takes_twice_its_products(prec):
    """BiCGStab(1) is BiCGStab (same iterations, same iterate).  IDR(1) with
    the shadow vector r_0 and minimal-residual omega generates BiCGStab's
    residuals every second operator application (Sonneveld & van Gijzen
    2008); with a random shadow vector and the angle-corrected omega used here
    its count of applications stays close to twice BiCGStab's iterations."""
    rp, col, val, rhs, _ = _mhfem_system(prec)
    x0 = np.zeros(rhs.size)
    xb, _, itb, _ = pyoracle.krylov(rp, col, val, rhs, x0, "bicgstab", prec, 1e-10)
    xl, cl, itl, _ = pyoracle.krylov(rp, col, val, rhs, x0, "bicgstabl", prec, 1e-10, restart=1)
    assert cl == 1 and abs(itl - itb) <= 1
    assert np.linalg.norm(xl - xb) / np.linalg.norm(xb) < 1e-8
    xi, ci, iti, _ = pyoracle.krylov(rp, col, val, rhs, x0, "idrs", prec, 1e-10, restart=1)
    assert ci == 1 and abs(iti - 2 * itb) <= 0.25 * 2 * itb, (iti, itb)


@pytest.mark.parametrize("solver,m", [("bicgstabl", 2), ("bicgstabl", 4), ("idrs", 2), ("idrs", 4), ("idrs", 8)])
@pytest.mark.parametrize("prec", ["none", "diagonal", "ilu0"])
def test_bicgstabl_and_idrs_against_direct_solve(solver, m, prec):
    rp, col, val, rhs, _ = _mhfem_system(prec)
    n = rp.size - 1
    xd = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(n, n)).tocsc(), rhs)
    x, conv, it, res = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), solver, prec, 1e-12, restart=m)
    assert conv == 1 and res < 1e-12
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_idrs_finite_termination():
    """In exact arithmetic IDR(s) solves an N x N system in at most N + N/s
    operator applications (Sonneveld & van Gijzen 2008); on a small
    well-conditioned nonsymmetric system the count stays near that bound."""
    rng = np.random.default_rng(8)
    N, s = 40, 4
    A = sp.csr_matrix(np.eye(N) * 4 + rng.uniform(-1, 1, (N, N)) * 0.3)
    A.sort_indices()
    b = rng.standard_normal(N)
    x, conv, it, res = pyoracle.krylov(A.indptr, A.indices, A.data, b, np.zeros(N), "idrs", "none", 1e-12, restart=s)
    assert conv == 1
    assert it <= N + N // s + 2, it
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-10


def test_sor_by_hand_and_against_direct_solve():
    """One SOR sweep on a 3x3 system worked by hand (w = 1.5, Gauss-Seidel
    order), and convergence to the direct solution on the MHFEM system for
    w = 1 (Gauss-Seidel) and w = 1.3."""
    A = sp.csr_matrix(np.array([[4.0, -1.0, 0.0], [-1.0, 4.0, -1.0], [0.0, -1.0, 4.0]]))
    b = np.array([3.0, 2.0, 3.0])
    x, conv, it, res = pyoracle.krylov(A.indptr, A.indices, A.data, b, np.zeros(3), "sor", "none", 1e-30,
                                       max_iter=1, restart=150)
    w = 1.5
    x0 = w * 3.0 / 4
    x1 = w * (2.0 + x0) / 4
    x2 = w * (3.0 + x1) / 4
    assert it == 1 and np.allclose(x, [x0, x1, x2], rtol=1e-15)
    rp, col, val, rhs, _ = _mhfem_system("none")
    n = rp.size - 1
    xd = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(n, n)).tocsc(), rhs)
    for m in (100, 130):
        x, conv, it, res = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), "sor", "none", 1e-11, max_iter=100000,
                                           restart=m)
        assert conv == 1
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8


def _ilut_dense(rp, col, val, n):
    frp, fcol, fval = pyoracle.ilut(rp, col, val, *_ILUT)
    F = sp.csr_matrix((fval, fcol, frp), shape=(n, n)).toarray()
    return np.tril(F, -1) + np.eye(n), np.triu(F), frp


_ILUT = (10, 1e-4)


def test_ilut_without_dropping_is_the_complete_lu():
    """tau = 0 and p >= N: nothing is dropped, so L U = A (the complete LU with
    fill-in), checked against the closed form on a tridiagonal matrix and by
    the product on an MHFEM system; BiCGStab then converges at once."""
    global _ILUT
    A = sp.csr_matrix(np.array([[4.0, 1.0, 0.0], [1.0, 4.0, 1.0], [0.0, 1.0, 4.0]]))
    _ILUT = (3, 0.0)
    L, U, _ = _ilut_dense(A.indptr, A.indices, A.data, 3)
    assert np.allclose(L, [[1, 0, 0], [0.25, 1, 0], [0, 1 / 3.75, 1]], rtol=1e-15)
    assert np.allclose(U, [[4, 1, 0], [0, 3.75, 1], [0, 0, 4 - 1 / 3.75]], rtol=1e-15)
    pr = problems.darcy_3d(s=2)
    o = pyoracle.Oracle(pr)
    o.pre_iterate(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    rp, col, val, rhs = [np.array(a) for a in o.csr()]
    n = rp.size - 1
    _ILUT = (n, 0.0)
    L, U, frp = _ilut_dense(rp, col, val, n)
    Ad = sp.csr_matrix((val, col, rp), shape=(n, n)).toarray()
    assert np.abs(L @ U - Ad).max() <= 1e-12 * np.abs(Ad).max()
    assert frp[-1] > rp[-1]  # fill-in was kept
    pyoracle.set_ilut(n, 0.0)
    x, conv, it, res = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), "bicgstab", "ilut", 1e-10)
    assert conv == 1 and it <= 1
    _ILUT = (10, 1e-4)
    pyoracle.set_ilut(10, 1e-4)


def test_ilut_with_p0_is_the_diagonal_and_default_beats_ilu0():
    global _ILUT
    rp, col, val, rhs, _ = _mhfem_system("none")
    n = rp.size - 1
    _ILUT = (0, 1e-4)
    L, U, _ = _ilut_dense(rp, col, val, n)
    Ad = sp.csr_matrix((val, col, rp), shape=(n, n)).toarray()
    assert np.array_equal(L, np.eye(n)) and np.array_equal(U, np.diag(np.diag(Ad)))
    _ILUT = (10, 1e-4)
    pyoracle.set_ilut(10, 1e-4)
    xd = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(n, n)).tocsc(), rhs)
    x, conv, it, res = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), "bicgstab", "ilut", 1e-12)
    _, _, it0, _ = pyoracle.krylov(rp, col, val, rhs, np.zeros(n), "bicgstab", "ilu0", 1e-12)
    assert conv == 1 and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert it <= it0, (it, it0)
