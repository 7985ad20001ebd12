"""Structured meshes, face numbering and seeded random fields.

These generate the synthetic inputs the benchmark configurations are defined
on (the reference reads VTK meshes through TNL; its example meshes are Git-LFS
stubs).  Vertex order inside each cell is fixed here because the scheme
depends on it: local face ``e`` of a simplex is the face opposite local
vertex ``e`` (reference MassMatrix.h:538,661,787,932); the faces of an edge
are its vertices, those of a quadrangle its edges (0,1), (1,2), (2,3), (3,0),
with vertices counter-clockwise from the lower-left one, and those of a
hexahedron (VTK vertex order) (0,1,2,3), (0,1,5,4), (1,2,6,5), (2,3,7,6),
(3,0,4,7), (4,5,6,7).
"""
from __future__ import annotations

import itertools

import numpy as np


def unit_square_triangles(s: int):
    """Unit square, ``s x s`` squares, each split by its (0,0)-(1,1) diagonal.

    Returns ``(points[nv,2], cells[nc,3])``.  Vertex ``(ix, iy)`` has index
    ``iy*(s+1)+ix``; squares are visited with ``ix`` fastest and contribute the
    triangles ``(v00, v10, v11)`` and ``(v00, v11, v01)``.
    """
    n1 = s + 1
    iy, ix = np.meshgrid(np.arange(n1), np.arange(n1), indexing="ij")
    points = np.stack([ix.ravel() / s, iy.ravel() / s], axis=1).astype(np.float64)
    qy, qx = np.meshgrid(np.arange(s), np.arange(s), indexing="ij")
    v00 = (qy * n1 + qx).ravel()
    v10 = v00 + 1
    v01 = v00 + n1
    v11 = v01 + 1
    t0 = np.stack([v00, v10, v11], axis=1)
    t1 = np.stack([v00, v11, v01], axis=1)
    cells = np.stack([t0, t1], axis=1).reshape(-1, 3).astype(np.int32)
    return points, cells


def unit_square_quads(s: int):
    """Unit square, ``s x s`` axis-aligned squares as quadrangle cells.

    Vertex ``(ix, iy)`` has index ``iy*(s+1)+ix``; cells are visited with
    ``ix`` fastest, vertices ``(v00, v10, v11, v01)`` (VTK_QUAD order).
    """
    n1 = s + 1
    iy, ix = np.meshgrid(np.arange(n1), np.arange(n1), indexing="ij")
    points = np.stack([ix.ravel() / s, iy.ravel() / s], axis=1).astype(np.float64)
    qy, qx = np.meshgrid(np.arange(s), np.arange(s), indexing="ij")
    v00 = (qy * n1 + qx).ravel()
    cells = np.stack([v00, v00 + 1, v00 + 1 + n1, v00 + n1], axis=1).astype(np.int32)
    return points, cells


def unit_interval(s: int):
    """Unit interval, ``s`` edges ``(k, k+1)``; points of shape ``(s+1, 1)``."""
    points = (np.arange(s + 1, dtype=np.float64) / s)[:, None]
    cells = np.stack([np.arange(s), np.arange(1, s + 1)], axis=1).astype(np.int32)
    return points, cells


def unit_cube_hexes(s: int):
    """Unit cube, ``s^3`` cubes as hexahedron cells.

    Vertex ``(ix, iy, iz)`` has index ``(iz*(s+1)+iy)*(s+1)+ix``; cells are
    visited with ``ix`` fastest, vertices in VTK_HEXAHEDRON order: the bottom
    face counter-clockwise from the lower-left corner, then the top face.
    """
    n1 = s + 1
    iz, iy, ix = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    points = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64) / s
    qz, qy, qx = np.meshgrid(np.arange(s), np.arange(s), np.arange(s), indexing="ij")
    v = ((qz * n1 + qy) * n1 + qx).ravel()
    up = n1 * n1
    bottom = [v, v + 1, v + 1 + n1, v + n1]
    cells = np.stack(bottom + [b + up for b in bottom], axis=1).astype(np.int32)
    return points, cells


def is_box(dim: int, vertices_per_cell: int) -> bool:
    """Edges, quadrangles and hexahedra are box cells; triangles and tetrahedra simplices."""
    return dim == 1 or vertices_per_cell == 2 ** dim


def faces_per_cell(dim: int, vertices_per_cell: int) -> int:
    return 2 * dim if is_box(dim, vertices_per_cell) else dim + 1


def face_vertex_table(dim: int, vertices_per_cell: int):
    """Local vertex positions of every local face."""
    if dim == 1:
        return [[0], [1]]
    if is_box(dim, vertices_per_cell) and dim == 2:
        return [[0, 1], [1, 2], [2, 3], [3, 0]]
    if is_box(dim, vertices_per_cell):
        return [[0, 1, 2, 3], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7], [4, 5, 6, 7]]
    return [[v for v in range(dim + 1) if v != e] for e in range(dim + 1)]


KUHN_PERMUTATIONS = list(itertools.permutations(range(3)))


def unit_cube_kuhn(s: int):
    """Unit cube, ``s^3`` cubes, each split into the 6 Kuhn tetrahedra.

    Vertex ``(ix, iy, iz)`` has index ``(iz*(s+1)+iy)*(s+1)+ix``; cubes are
    visited with ``ix`` fastest.  Tetrahedron ``k`` of a cube follows the
    monotone path ``v000 -> +e_a -> +e_a+e_b -> v111`` for the k-th permutation
    ``(a, b, c)`` of the axes in lexicographic order.  All cubes share the same
    diagonal direction, so the triangulation is conforming.
    """
    n1 = s + 1
    iz, iy, ix = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    points = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64) / s
    qz, qy, qx = np.meshgrid(np.arange(s), np.arange(s), np.arange(s), indexing="ij")
    base = ((qz * n1 + qy) * n1 + qx).ravel()
    step = np.array([1, n1, n1 * n1])
    tets = []
    for perm in KUHN_PERMUTATIONS:
        a, b, _ = perm
        v0 = base
        v1 = base + step[a]
        v2 = base + step[a] + step[b]
        v3 = base + step.sum()
        tets.append(np.stack([v0, v1, v2, v3], axis=1))
    cells = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return points, cells


def face_numbering(cells: np.ndarray, num_points: int, dim: int = None):
    """First-appearance face numbering.

    Faces are numbered in the order in which they are first met when cells are
    visited in index order and the faces of a cell in local order (see the
    module docstring for the local faces); the cells of a face are in
    ascending order, the second one is -1 on boundary faces.  ``dim`` is the
    mesh dimension (default: a simplex mesh, ``vertices - 1``).
    Returns ``(cell_faces[nc,fpc] int32, face_cells[nf,2] int32)``.
    """
    nc, nvc = cells.shape
    dim = nvc - 1 if dim is None else dim
    table = face_vertex_table(dim, nvc)
    fpc, d = len(table), len(table[0])
    seeds = np.empty((nc, fpc, d), dtype=np.int64)
    for e, verts in enumerate(table):
        seeds[:, e, :] = cells[:, verts]
    seeds.sort(axis=2)
    bits = int(np.ceil(np.log2(max(num_points, 2))))
    if bits * d > 63:
        raise ValueError("mesh too large for 63-bit face keys")
    key = np.zeros((nc, fpc), dtype=np.int64)
    for t in range(d):
        key = (key << bits) | seeds[:, :, t]
    key = key.ravel()
    uniq, first, inverse = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    cell_faces = rank[inverse].reshape(nc, fpc).astype(np.int32)
    nf = uniq.size
    flat = cell_faces.ravel()
    owner = np.repeat(np.arange(nc, dtype=np.int64), fpc)
    face_cells = np.full((nf, 2), -1, dtype=np.int64)
    # visiting (K, e) in ascending order: the first visit of a face is its first cell
    srt = np.lexsort((owner, flat))
    f_sorted = flat[srt]
    k_sorted = owner[srt]
    start = np.ones(f_sorted.size, dtype=bool)
    start[1:] = f_sorted[1:] != f_sorted[:-1]
    face_cells[f_sorted[start], 0] = k_sorted[start]
    second = ~start
    if np.any(second[1:] & second[:-1]):
        raise ValueError("non-conforming mesh: a face has more than two cells")
    face_cells[f_sorted[second], 1] = k_sorted[second]
    return cell_faces, face_cells.astype(np.int32)


def face_centroids(points, cells, cell_faces, nf):
    """Centroid of every face (from the first cell that contains it)."""
    dim = points.shape[1]
    cent = np.zeros((nf, dim))
    for e, verts in enumerate(face_vertex_table(dim, cells.shape[1])):
        cent[cell_faces[:, e]] = points[cells[:, verts]].mean(axis=1)
    return cent


def cell_centroids(points, cells):
    return points[cells].mean(axis=1)


def outward_face_vectors(points, cells):
    """Area-weighted outward normal of every local face: ``|E_e| n_{K,e}``.

    Shape ``(nc, fpc, dim)``.  Used to project a constant vector field onto the
    lowest-order Raviart-Thomas degrees of freedom (face fluxes).
    """
    nc, nvc = cells.shape
    dim = points.shape[1]
    P = points[cells]  # (nc, nvc, dim)
    if is_box(dim, nvc):
        # |E| n on the faces of an axis-aligned box: -y, +x, +y, -x (an edge: -x, +x)
        if dim == 1:
            return np.stack([-np.ones(nc), np.ones(nc)], axis=1)[:, :, None]
        if dim == 3:
            # faces -z, -y, +x, +y, -x, +z
            h = P[:, 6] - P[:, 0]
            out = np.zeros((nc, 6, 3))
            for e, (a, sign) in enumerate([(2, -1), (1, -1), (0, 1), (1, 1), (0, -1), (2, 1)]):
                out[:, e, a] = sign * np.prod(np.delete(h, a, axis=1), axis=1)
            return out
        hx = P[:, 2, 0] - P[:, 0, 0]
        hy = P[:, 2, 1] - P[:, 0, 1]
        out = np.zeros((nc, 4, 2))
        out[:, 0, 1] = -hx
        out[:, 1, 0] = hy
        out[:, 2, 1] = hx
        out[:, 3, 0] = -hy
        return out
    fpc = nvc
    out = np.zeros((nc, fpc, dim))
    for e in range(fpc):
        others = [v for v in range(fpc) if v != e]
        if dim == 2:
            t = P[:, others[1]] - P[:, others[0]]
            nvec = np.stack([t[:, 1], -t[:, 0]], axis=1)
        else:
            nvec = 0.5 * np.cross(P[:, others[1]] - P[:, others[0]], P[:, others[2]] - P[:, others[0]])
        to_opp = P[:, e] - P[:, others[0]]
        sign = np.where((nvec * to_opp).sum(axis=1) > 0, -1.0, 1.0)
        out[:, e] = nvec * sign[:, None]
    return out


def gaussian_field(shape, sigma, corr_len, seed, spacing=None):
    """Seeded stationary Gaussian random field, exponential covariance.

    ``C(r) = sigma^2 exp(-r / corr_len)``, generated by circulant embedding on
    a periodic grid twice the size in each direction (negative eigenvalues of
    the embedding are clipped), then cropped.  ``spacing`` defaults to
    ``1/shape`` (unit domain, one value per cell of the structured grid).
    """
    shape = tuple(int(s) for s in shape)
    if spacing is None:
        spacing = tuple(1.0 / s for s in shape)
    big = tuple(2 * s for s in shape)
    axes = []
    for s, h in zip(big, spacing):
        k = np.arange(s)
        k = np.minimum(k, s - k)
        axes.append(k * h)
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum(g * g for g in grids))
    cov = sigma * sigma * np.exp(-r / corr_len)
    lam = np.fft.fftn(cov).real
    lam = np.clip(lam, 0.0, None)
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(big) + 1j * rng.standard_normal(big)
    field = np.fft.ifftn(np.sqrt(lam) * noise) * np.sqrt(np.prod(big))
    field = field.real
    sl = tuple(slice(0, s) for s in shape)
    return np.ascontiguousarray(field[sl])
