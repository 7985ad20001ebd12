"""The benchmark configurations as model-coefficient sets.

A :class:`Problem` carries everything the reference takes from its mesh file,
boundary-conditions file, initial condition and model ``init``:
the mesh, the number of equations ``n``, the per-element coefficients
``N, u, m, D, w, a, r, f`` in the reference's CUDA layouts
(BaseModel.h:50-124; a 0-d array means a uniform coefficient,
UniformCoefficient.h:7-53), the boundary tags/values/Dirichlet values in the
``i*F+E`` layout (BoundaryConditionsStorage.h:9-15), the initial ``Z_iK`` and
the time-stepping and solver settings.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import meshgen

FIXED_VALUE, FIXED_FLUX, ADVECTIVE_OUTFLOW, FIXED_FLUX_NO_ADVECTION = 0, 1, 2, 3
EXPLICIT_UPWIND, IMPLICIT_UPWIND, IMPLICIT_TRACE = 0, 1, 2

COEFFICIENTS = ("N_ijK", "u_ijKe", "m_iK", "D_ijK", "w_iKe", "a_ijKe", "r_ijK", "f_iK")


@dataclass
class Problem:
    name: str
    points: np.ndarray
    cells: np.ndarray
    n: int
    coeffs: dict
    tags: np.ndarray
    bc_values: np.ndarray
    bc_dirichlet: np.ndarray
    Z_iK0: np.ndarray
    dt: float
    steps: int
    solver: str = "bicgstab"
    preconditioner: str = "diagonal"
    tol: float = 1e-10
    lumping: int = 0
    advection: int = EXPLICIT_UPWIND
    do_mobility_upwind: int = 1
    bc_mobility: Optional[np.ndarray] = None
    cell_faces: Optional[np.ndarray] = None
    face_cells: Optional[np.ndarray] = None
    # time-dependent Dirichlet values: f(time) -> array (i*F+E layout) or None
    dirichlet_at: Optional[Callable[[float], np.ndarray]] = None
    # the same boundary model as device time functions (mhfem_time_function):
    # {("dirichlet" | "neumann", species): dict(kind=..., ...)}; the stored
    # values times f(time + tau).  When set, the library evaluates the model in
    # the assembly; dirichlet_at then gives the same values on the host.
    boundary_functions: dict = field(default_factory=dict)
    extra: dict = field(default_factory=dict)

    @property
    def dim(self):
        return self.points.shape[1]

    @property
    def num_cells(self):
        return self.cells.shape[0]

    @property
    def num_faces(self):
        return self.face_cells.shape[0]

    @property
    def box(self):
        """Edge, quadrangle or hexahedron cells (else triangles or tetrahedra)."""
        return meshgen.is_box(self.dim, self.cells.shape[1])

    @property
    def fpc(self):
        return meshgen.faces_per_cell(self.dim, self.cells.shape[1])

    def coefficient_array(self, name):
        """Full (non-uniform) array of a coefficient in the CUDA layout."""
        n, C, fpc = self.n, self.num_cells, self.fpc
        size = {
            "N_ijK": n * n * C, "D_ijK": n * n * C, "r_ijK": n * n * C,
            "u_ijKe": n * n * fpc * C, "a_ijKe": n * n * fpc * C,
            "m_iK": n * C, "f_iK": n * C, "w_iKe": n * fpc * C,
        }[name]
        v = np.asarray(self.coeffs[name], dtype=np.float64)
        if v.ndim == 0:
            return np.full(size, float(v))
        assert v.size == size, (name, v.size, size)
        return v.ravel()


def _topology(points, cells):
    cf, fc = meshgen.face_numbering(cells, points.shape[0], points.shape[1])
    return cf, fc


def _boundary_mask(fc):
    return fc[:, 1] < 0


def _ij_blocks(n, C, blocks):
    """Stack per-(i,j) cell arrays into the (i,j,K) layout."""
    out = np.empty((n, n, C))
    for i in range(n):
        for j in range(n):
            out[i, j] = blocks[i][j]
    return out.ravel()


def heat_2d(s: int = 64, dt: float = 1e-4, steps: int = 100, lumping: int = 0) -> Problem:
    """2-D heat equation on the unit square (HeatEquationModel.h:114-133):
    N=1, m=1, D=1, u=w=a=r=f=0, zero Dirichlet on the whole boundary
    (init.py:38-40), Z0 = exp(-((x-.5)^2+(y-.5)^2)/0.01) at cell centroids,
    BiCGStab + Jacobi to 1e-10."""
    points, cells = meshgen.unit_square_triangles(s)
    cf, fc = _topology(points, cells)
    F = fc.shape[0]
    coeffs = dict(N_ijK=np.array(1.0), u_ijKe=np.array(0.0), m_iK=np.array(1.0),
                  D_ijK=np.array(1.0), w_iKe=np.array(0.0), a_ijKe=np.array(0.0),
                  r_ijK=np.array(0.0), f_iK=np.array(0.0))
    c = meshgen.cell_centroids(points, cells)
    z0 = np.exp(-((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.01)
    return Problem("heat_2d", points, cells, 1, coeffs,
                   tags=np.full(F, FIXED_VALUE, np.uint8), bc_values=np.zeros(F),
                   bc_dirichlet=np.zeros(F), Z_iK0=z0, dt=dt, steps=steps,
                   solver="bicgstab", preconditioner="diagonal", lumping=lumping,
                   do_mobility_upwind=0, bc_mobility=np.ones(F), cell_faces=cf, face_cells=fc)


def _cube_lnK(s, sigma, corr_len, seed):
    g = meshgen.gaussian_field((s, s, s), sigma, corr_len, seed)  # indexed [ix, iy, iz]
    # cube (ix,iy,iz) is visited with ix fastest -> transpose to [iz,iy,ix]
    per_cube = np.transpose(g, (2, 1, 0)).ravel()
    return np.repeat(per_cube, 6)


def darcy_3d(s: int = 64, seed: int = 1, dt: float = 1e-3, steps: int = 20,
             preconditioner: str = "ilu0", shuffle_cells: Optional[int] = None) -> Problem:
    """3-D Darcy-type pressure on the unit cube, Kuhn tetrahedra: n=1, N=0.2,
    m=1, D=K with ln K a seeded Gaussian field (exponential covariance,
    sigma=1, correlation length 0.1, constant per cube); Z=1 at x=0, Z=0 at
    x=1, no-flow (FixedFluxNoAdvection, J=0) elsewhere; Z0=0.
    shuffle_cells: the cells (and with them the face numbering, faces being
    numbered in order of first appearance) in a random order of that seed."""
    points, cells = meshgen.unit_cube_kuhn(s)
    K = np.exp(_cube_lnK(s, 1.0, 0.1, seed))
    if shuffle_cells is not None:
        perm = np.random.default_rng(shuffle_cells).permutation(cells.shape[0])
        cells, K = np.ascontiguousarray(cells[perm]), K[perm]
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    coeffs = dict(N_ijK=np.array(0.2), u_ijKe=np.array(0.0), m_iK=np.array(1.0), D_ijK=K,
                  w_iKe=np.array(0.0), a_ijKe=np.array(0.0), r_ijK=np.array(0.0),
                  f_iK=np.array(0.0))
    fcen = meshgen.face_centroids(points, cells, cf, F)
    bnd = _boundary_mask(fc)
    tags = np.full(F, FIXED_VALUE, np.uint8)
    tags[bnd] = FIXED_FLUX_NO_ADVECTION
    x0 = bnd & (np.abs(fcen[:, 0]) < 1e-12)
    x1 = bnd & (np.abs(fcen[:, 0] - 1.0) < 1e-12)
    tags[x0 | x1] = FIXED_VALUE
    dirichlet = np.zeros(F)
    dirichlet[x0] = 1.0
    return Problem("darcy_3d", points, cells, 1, coeffs, tags=tags, bc_values=np.zeros(F),
                   bc_dirichlet=dirichlet, Z_iK0=np.zeros(C), dt=dt, steps=steps,
                   solver="bicgstab", preconditioner=preconditioner, do_mobility_upwind=1,
                   bc_mobility=np.ones(F), cell_faces=cf, face_cells=fc,
                   extra=dict(seed=seed, s=s))


def _adr_coefficients(n, C, K, flux):
    """Coupled ADR coefficients of n species: D_ij = K (i = j) or 0.1 K,
    N = 0.2 I, m = 1, r_ij = 1e-2 (i = j) or -5e-3, a_iiKe = flux, u = w = f = 0."""
    zero = np.zeros(C)
    D = _ij_blocks(n, C, [[K if i == j else 0.1 * K for j in range(n)] for i in range(n)])
    N = _ij_blocks(n, C, [[0.2 + zero if i == j else zero for j in range(n)] for i in range(n)])
    r = _ij_blocks(n, C, [[(1e-2 if i == j else -5e-3) + zero for j in range(n)] for i in range(n)])
    a = np.zeros((n, n, flux.shape[1], C))
    for i in range(n):
        a[i, i] = flux.T
    return dict(N_ijK=N, u_ijKe=np.array(0.0), m_iK=np.array(1.0), D_ijK=D,
                w_iKe=np.array(0.0), a_ijKe=a.ravel(), r_ijK=r, f_iK=np.array(0.0))


def _adr_boundary(n, fc, fcen):
    """Dirichlet Z_i = 2^-i at x = 0, AdvectiveOutflow at x = 1,
    FixedFluxNoAdvection (J = 0) elsewhere; interior tags are unused."""
    F = fc.shape[0]
    bnd = _boundary_mask(fc)
    x0 = bnd & (np.abs(fcen[:, 0]) < 1e-12)
    x1 = bnd & (np.abs(fcen[:, 0] - 1.0) < 1e-12)
    tags = np.full((n, F), FIXED_VALUE, np.uint8)
    tags[:, bnd] = FIXED_FLUX_NO_ADVECTION
    tags[:, x1] = ADVECTIVE_OUTFLOW
    tags[:, x0] = FIXED_VALUE
    dirichlet = np.zeros((n, F))
    for i in range(n):
        dirichlet[i, x0] = 0.5 ** i
    return tags, dirichlet, x0


def adr_3d(s: int = 64, seed: int = 1, dt: float = 1e-3, steps: int = 50,
           preconditioner: str = "ilu0", time_dependent: bool = False, n: int = 2) -> Problem:
    """n coupled species (default 2) of advection-diffusion-reaction on the
    Darcy mesh and K field: D11=D22=K, D12=D21=0.1K, N=0.2 I, m=1, constant
    advection a_11=a_22=(1,.5,.25) projected onto the RTN0 face fluxes
    (a_iiKe = |E_e| a.n_{K,e}), explicit upwind,
    r=[[1e-2,-5e-3],[-5e-3,1e-2]], f=0, Z0=0; Dirichlet Z=(1,0.5) at x=0,
    AdvectiveOutflow at x=1, FixedFluxNoAdvection (J=0) elsewhere (for n != 2
    the same pattern: D_ij = 0.1 K and r_ij = -5e-3 off the diagonal,
    Z_i = 2^-i at x=0).  ``time_dependent`` makes Z_1 = sin(2 pi t / 0.5)
    at x=0 (evaluated at time+tau, Solver.hpp:948)."""
    points, cells = meshgen.unit_cube_kuhn(s)
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    K = np.exp(_cube_lnK(s, 1.0, 0.1, seed))
    fv = meshgen.outward_face_vectors(points, cells)  # (C, 4, 3)
    flux = (fv * np.array([1.0, 0.5, 0.25])).sum(axis=2)  # (C, 4)
    coeffs = _adr_coefficients(n, C, K, flux)
    fcen = meshgen.face_centroids(points, cells, cf, F)
    tags, dirichlet, x0 = _adr_boundary(n, fc, fcen)
    dirichlet_at = None
    functions = {}
    if time_dependent:
        # stored Z_1 = 1 at x = 0, times sin(2 pi t / 0.5)
        base = dirichlet.ravel().copy()
        functions[("dirichlet", 0)] = dict(kind="sine", amplitude=1.0, angular_frequency=2 * np.pi / 0.5)
        dirichlet_at = boundary_values_at(base, n, F, functions)
    return Problem("adr_3d", points, cells, n, coeffs, tags=tags.ravel(),
                   bc_values=np.zeros(n * F), bc_dirichlet=dirichlet.ravel(),
                   Z_iK0=np.zeros(n * C), dt=dt, steps=steps, solver="bicgstab",
                   preconditioner=preconditioner, do_mobility_upwind=1,
                   bc_mobility=np.ones(n * F), cell_faces=cf, face_cells=fc,
                   dirichlet_at=dirichlet_at, boundary_functions=functions, extra=dict(seed=seed, s=s))


def time_function(spec, t):
    """A boundary time function (the library's mhfem_time_function) at t,
    on the host."""
    kind = spec["kind"]
    if kind == "constant":
        return spec.get("offset", 0.0)
    if kind == "sine":
        return spec.get("offset", 0.0) + spec.get("amplitude", 0.0) * np.sin(
            spec.get("angular_frequency", 0.0) * t + spec.get("phase", 0.0))
    if kind == "table":
        return float(np.interp(t, spec["times"], spec["values"]))
    raise ValueError(kind)


def boundary_values_at(base, n, F, functions, which="dirichlet"):
    """The values of a separable boundary model at time t as an array in the
    i*F+E layout: the stored values of species i times its time function."""
    base = np.asarray(base, np.float64).ravel()

    def values_at(t):
        out = base.copy()
        for (w, i), spec in functions.items():
            if w == which:
                out[i * F:(i + 1) * F] *= time_function(spec, t)
        return out

    return values_at


def adr_2d(s: int = 16, n: int = 2, dt: float = 1e-3, steps: int = 3,
           preconditioner: str = "diagonal") -> Problem:
    """The coupled ADR model of adr_3d on the triangulated unit square:
    K = 1 + 0.5 sin(2 pi x) sin(2 pi y) at cell centroids, advection
    (1, 0.5) projected onto the face fluxes, the same boundary pattern."""
    points, cells = meshgen.unit_square_triangles(s)
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    c = meshgen.cell_centroids(points, cells)
    K = 1.0 + 0.5 * np.sin(2 * np.pi * c[:, 0]) * np.sin(2 * np.pi * c[:, 1])
    fv = meshgen.outward_face_vectors(points, cells)  # (C, 3, 2)
    flux = (fv * np.array([1.0, 0.5])).sum(axis=2)
    coeffs = _adr_coefficients(n, C, K, flux)
    fcen = meshgen.face_centroids(points, cells, cf, F)
    tags, dirichlet, _ = _adr_boundary(n, fc, fcen)
    return Problem("adr_2d", points, cells, n, coeffs, tags=tags.ravel(),
                   bc_values=np.zeros(n * F), bc_dirichlet=dirichlet.ravel(),
                   Z_iK0=np.zeros(n * C), dt=dt, steps=steps, solver="bicgstab",
                   preconditioner=preconditioner, do_mobility_upwind=1,
                   bc_mobility=np.ones(n * F), cell_faces=cf, face_cells=fc)


def heat_quads(s: int = 16, dt: float = 1e-3, steps: int = 3, lumping: int = 0) -> Problem:
    """The heat problem of heat_2d on ``s x s`` square quadrangle cells."""
    points, cells = meshgen.unit_square_quads(s)
    cf, fc = _topology(points, cells)
    F = fc.shape[0]
    coeffs = dict(N_ijK=np.array(1.0), u_ijKe=np.array(0.0), m_iK=np.array(1.0),
                  D_ijK=np.array(1.0), w_iKe=np.array(0.0), a_ijKe=np.array(0.0),
                  r_ijK=np.array(0.0), f_iK=np.array(0.0))
    c = meshgen.cell_centroids(points, cells)
    z0 = np.exp(-((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.01)
    return Problem("heat_quads", points, cells, 1, coeffs,
                   tags=np.full(F, FIXED_VALUE, np.uint8), bc_values=np.zeros(F),
                   bc_dirichlet=np.zeros(F), Z_iK0=z0, dt=dt, steps=steps,
                   solver="bicgstab", preconditioner="diagonal", lumping=lumping,
                   do_mobility_upwind=0, bc_mobility=np.ones(F), cell_faces=cf, face_cells=fc)


def adr_quads(s: int = 16, n: int = 2, dt: float = 1e-3, steps: int = 3,
              preconditioner: str = "diagonal", lumping: int = 0) -> Problem:
    """The coupled ADR model of adr_2d on ``s x s`` square quadrangle cells
    (MassMatrix.h:146-303): the same K field, advection (1, 0.5) as face
    fluxes, the same boundary pattern."""
    points, cells = meshgen.unit_square_quads(s)
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    c = meshgen.cell_centroids(points, cells)
    K = 1.0 + 0.5 * np.sin(2 * np.pi * c[:, 0]) * np.sin(2 * np.pi * c[:, 1])
    fv = meshgen.outward_face_vectors(points, cells)  # (C, 4, 2)
    flux = (fv * np.array([1.0, 0.5])).sum(axis=2)
    coeffs = _adr_coefficients(n, C, K, flux)
    fcen = meshgen.face_centroids(points, cells, cf, F)
    tags, dirichlet, _ = _adr_boundary(n, fc, fcen)
    return Problem("adr_quads", points, cells, n, coeffs, tags=tags.ravel(),
                   bc_values=np.zeros(n * F), bc_dirichlet=dirichlet.ravel(),
                   Z_iK0=np.zeros(n * C), dt=dt, steps=steps, solver="bicgstab",
                   preconditioner=preconditioner, lumping=lumping, do_mobility_upwind=1,
                   bc_mobility=np.ones(n * F), cell_faces=cf, face_cells=fc)


def adr_hexes(s: int = 6, n: int = 2, dt: float = 1e-3, steps: int = 3,
              preconditioner: str = "diagonal", lumping: int = 0, seed: int = 1) -> Problem:
    """The coupled ADR model of adr_3d on ``s^3`` cube (voxel) cells
    (MassMatrix.h:305-508): ln K the seeded Gaussian field, advection
    (1, .5, .25) as face fluxes, the same boundary pattern."""
    points, cells = meshgen.unit_cube_hexes(s)
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    g = meshgen.gaussian_field((s, s, s), 1.0, 0.1, seed)  # [ix, iy, iz]
    K = np.exp(np.transpose(g, (2, 1, 0)).ravel())
    fv = meshgen.outward_face_vectors(points, cells)  # (C, 6, 3)
    flux = (fv * np.array([1.0, 0.5, 0.25])).sum(axis=2)
    coeffs = _adr_coefficients(n, C, K, flux)
    fcen = meshgen.face_centroids(points, cells, cf, F)
    tags, dirichlet, _ = _adr_boundary(n, fc, fcen)
    return Problem("adr_hexes", points, cells, n, coeffs, tags=tags.ravel(),
                   bc_values=np.zeros(n * F), bc_dirichlet=dirichlet.ravel(),
                   Z_iK0=np.zeros(n * C), dt=dt, steps=steps, solver="bicgstab",
                   preconditioner=preconditioner, lumping=lumping, do_mobility_upwind=1,
                   bc_mobility=np.ones(n * F), cell_faces=cf, face_cells=fc)


def adr_1d(s: int = 32, n: int = 2, dt: float = 1e-3, steps: int = 3,
           preconditioner: str = "diagonal", lumping: int = 0) -> Problem:
    """The coupled ADR model on the unit interval (edge cells, MassMatrix.h:28-145):
    K = 1 + 0.5 sin(2 pi x), advection 1 as face fluxes (+-1), Dirichlet
    Z_i = 2^-i at x = 0, AdvectiveOutflow at x = 1."""
    points, cells = meshgen.unit_interval(s)
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    c = meshgen.cell_centroids(points, cells)
    K = 1.0 + 0.5 * np.sin(2 * np.pi * c[:, 0])
    flux = meshgen.outward_face_vectors(points, cells)[:, :, 0]  # (C, 2)
    coeffs = _adr_coefficients(n, C, K, flux)
    fcen = meshgen.face_centroids(points, cells, cf, F)
    tags, dirichlet, _ = _adr_boundary(n, fc, fcen)
    return Problem("adr_1d", points, cells, n, coeffs, tags=tags.ravel(),
                   bc_values=np.zeros(n * F), bc_dirichlet=dirichlet.ravel(),
                   Z_iK0=np.zeros(n * C), dt=dt, steps=steps, solver="bicgstab",
                   preconditioner=preconditioner, lumping=lumping, do_mobility_upwind=1,
                   bc_mobility=np.ones(n * F), cell_faces=cf, face_cells=fc)


def from_files(mesh: str, boundary_conditions_file: Optional[str] = None,
               initial_condition: Optional[str] = None, n: int = 1, coeffs: Optional[dict] = None,
               dt: float = 1e-3, steps: int = 1, solver: str = "bicgstab",
               preconditioner: str = "diagonal", tol: float = 1e-10, lumping: int = 0,
               advection: int = EXPLICIT_UPWIND, do_mobility_upwind: int = 1,
               mesh_format: str = "auto") -> Problem:
    """A problem from the reference's input files, as control.h::init and
    Solver::setInitialCondition read them (control.h:20-48,
    Solver.hpp:262-309): the mesh (.vtu or .vtk), the boundary conditions
    (BoundaryConditionsStorage in TNL's format; zero Dirichlet everywhere when
    absent) and the "InitialCondition[Z<i>]" cell arrays (Z_iK = 0 when
    absent).  ``coeffs`` are the model's coefficients (default: the heat
    model with unit diffusivity)."""
    from . import capi

    mf = capi.MeshFile(mesh, mesh_format)
    points, cells = mf.points, mf.cells
    cf, fc = _topology(points, cells)
    F, C = fc.shape[0], cells.shape[0]
    dofs = n * F
    if boundary_conditions_file is not None:
        bcs = capi.BoundaryConditionsStorage.load(boundary_conditions_file)
        if bcs.dofSize != dofs:
            raise ValueError(f"Wrong dofSize in BoundaryConditionsStorage loaded from file "
                             f"{boundary_conditions_file}. Expected {dofs} elements, got {bcs.dofSize}.")
        tags, values, dirichlet = bcs.tags, bcs.values, bcs.dirichletValues
    else:
        tags = np.full(dofs, FIXED_VALUE, np.uint8)
        values, dirichlet = np.zeros(dofs), np.zeros(dofs)
    if initial_condition is not None:
        ic = capi.MeshFile(initial_condition)
        if ic.cells.shape[0] != C:
            raise ValueError(f"{initial_condition}: {ic.cells.shape[0]} cells, the mesh has {C}")
        Z0 = np.concatenate([ic.cell_data(f"InitialCondition[Z{i}]") for i in range(n)])
    else:
        Z0 = np.zeros(n * C)
    if coeffs is None:
        coeffs = dict(N_ijK=np.array(1.0), u_ijKe=np.array(0.0), m_iK=np.array(1.0),
                      D_ijK=np.array(1.0), w_iKe=np.array(0.0), a_ijKe=np.array(0.0),
                      r_ijK=np.array(0.0), f_iK=np.array(0.0))
    return Problem("from_files", points, cells, n, coeffs, tags=tags, bc_values=values,
                   bc_dirichlet=dirichlet, Z_iK0=Z0, dt=dt, steps=steps, solver=solver,
                   preconditioner=preconditioner, tol=tol, lumping=lumping, advection=advection,
                   do_mobility_upwind=do_mobility_upwind, bc_mobility=np.ones(dofs),
                   cell_faces=cf, face_cells=fc)


def ensemble_sizes(seed: int, count: int = 96, lo: int = 32, hi: int = 256):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count)


def ensemble_member(s: int, seed: int, dt: float = 1e-4, steps: int = 50) -> Problem:
    """One member of the batched ensemble: heat-type model on an s x s square
    mesh with D = K, ln K Gaussian (sigma=1.5, correlation length 0.05)."""
    points, cells = meshgen.unit_square_triangles(s)
    cf, fc = _topology(points, cells)
    F = fc.shape[0]
    g = meshgen.gaussian_field((s, s), 1.5, 0.05, seed)  # [ix, iy]
    per_sq = g.T.ravel()
    K = np.exp(np.repeat(per_sq, 2))
    coeffs = dict(N_ijK=np.array(1.0), u_ijKe=np.array(0.0), m_iK=np.array(1.0), D_ijK=K,
                  w_iKe=np.array(0.0), a_ijKe=np.array(0.0), r_ijK=np.array(0.0),
                  f_iK=np.array(0.0))
    c = meshgen.cell_centroids(points, cells)
    z0 = np.exp(-((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.01)
    return Problem(f"ensemble_{s}", points, cells, 1, coeffs,
                   tags=np.full(F, FIXED_VALUE, np.uint8), bc_values=np.zeros(F),
                   bc_dirichlet=np.zeros(F), Z_iK0=z0, dt=dt, steps=steps,
                   solver="bicgstab", preconditioner="diagonal", do_mobility_upwind=0,
                   bc_mobility=np.ones(F), cell_faces=cf, face_cells=fc, extra=dict(seed=seed))


def concatenate(members):
    """One problem over the disjoint union of the members' meshes (cells and
    faces numbered member by member), for a batched context: every array is
    the concatenation of the members' arrays along the cell (or face) index,
    which is the fastest index of every layout."""
    n = members[0].n
    fpc = members[0].fpc
    assert all(m.n == n and m.fpc == fpc for m in members)
    pts, cells, voff = [], [], 0
    for m in members:
        pts.append(m.points)
        cells.append(m.cells + voff)
        voff += m.points.shape[0]
    points = np.concatenate(pts)
    cells = np.concatenate(cells).astype(np.int32)
    cf, fc = meshgen.face_numbering(cells, points.shape[0])
    lead = {"N_ijK": (n, n), "D_ijK": (n, n), "r_ijK": (n, n), "u_ijKe": (n, n, fpc), "a_ijKe": (n, n, fpc),
            "m_iK": (n,), "f_iK": (n,), "w_iKe": (n, fpc)}
    coeffs = {}
    for name, shape in lead.items():
        vals = [np.asarray(m.coeffs[name], np.float64) for m in members]
        if all(v.ndim == 0 for v in vals) and all(float(v) == float(vals[0]) for v in vals):
            coeffs[name] = np.array(float(vals[0]))
        else:
            parts = [m.coefficient_array(name).reshape(shape + (m.num_cells,)) for m in members]
            coeffs[name] = np.concatenate(parts, axis=-1).ravel()

    def cat_faces(attr, dtype):
        return np.concatenate([np.asarray(getattr(m, attr), dtype).reshape(n, m.num_faces) for m in members],
                              axis=-1).ravel()

    z0 = np.concatenate([np.asarray(m.Z_iK0, np.float64).reshape(n, m.num_cells) for m in members], axis=-1).ravel()
    first = members[0]
    return Problem(f"batch_of_{len(members)}", points, cells, n, coeffs, tags=cat_faces("tags", np.uint8),
                   bc_values=cat_faces("bc_values", np.float64), bc_dirichlet=cat_faces("bc_dirichlet", np.float64),
                   Z_iK0=z0, dt=first.dt, steps=first.steps, solver=first.solver, preconditioner=first.preconditioner,
                   tol=first.tol, lumping=first.lumping, advection=first.advection,
                   do_mobility_upwind=first.do_mobility_upwind, bc_mobility=cat_faces("bc_mobility", np.float64),
                   cell_faces=cf, face_cells=fc,
                   extra=dict(instances=[(m.points, m.cells) for m in members],
                              cell_offsets=np.cumsum([0] + [m.num_cells for m in members]),
                              face_offsets=np.cumsum([0] + [m.num_faces for m in members])))


def ensemble(seed: int = 7, count: int = 96, lo: int = 32, hi: int = 256, steps: int = 50):
    """The batched ensemble: `count` members with sizes drawn from the seed."""
    sizes = ensemble_sizes(seed, count, lo, hi)
    return concatenate([ensemble_member(int(s), seed * 1000 + k, steps=steps) for k, s in enumerate(sizes)])
