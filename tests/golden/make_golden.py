"""Writes the regression fixtures of this directory from the C oracle.

The reference tree holds no golden vectors for this path, so these fixtures
are the oracle's own outputs, kept to catch unintended changes: the CSR
system after the first preIterate + assemble, and Z_iF / Z_iK after two time
steps solved to 1e-13, for a triangle mesh (n = 1) and a tetrahedron mesh
(n = 2, coupled, advection, all four boundary types).
Run from the repository root: python tests/golden/make_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tnl-mhfem_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from mhfem_amd import problems  # noqa: E402
import pyoracle  # noqa: E402


def cases():
    tri = problems.heat_2d(s=4, steps=2)
    tet = problems.adr_3d(s=2, steps=2, preconditioner="diagonal")
    # every boundary face cycles through the four tags, with fixed values
    n, F = tet.n, tet.num_faces
    tags = np.array(tet.tags, copy=True)
    vals = np.zeros(n * F)
    dirv = np.zeros(n * F)
    kinds = [problems.FIXED_VALUE, problems.FIXED_FLUX, problems.ADVECTIVE_OUTFLOW,
             problems.FIXED_FLUX_NO_ADVECTION]
    for i in range(n):
        for q, E in enumerate(np.nonzero(tet.face_cells[:, 1] < 0)[0]):
            tags[i * F + E] = kinds[(q + i) % 4]
            vals[i * F + E] = 0.25 * ((q % 7) - 3)
            dirv[i * F + E] = 0.5 + 0.1 * (q % 5)
    import dataclasses
    tet = dataclasses.replace(tet, tags=tags, bc_values=vals, bc_dirichlet=dirv)
    return {"heat2d_s4": tri, "adr3d_s2_tags": tet}


def compute(pr):
    o = pyoracle.Oracle(pr)
    o.pre_iterate(0.0, pr.dt)
    o.assemble(0.0, pr.dt)
    # the oracle's arrays are views of its own storage: copy before it moves on
    system = {k: v.copy() for k, v in zip(("rowptr", "col", "val", "rhs"), o.csr())}
    o.solve(tol=1e-13)
    o.post_iterate(0.0, pr.dt)
    o.step(pr.dt, pr.dt, tol=1e-13)
    return dict(system, Z_iF=o.array("Z_iF").copy(), Z_iK=o.array("Z_iK").copy())


if __name__ == "__main__":
    for name, pr in cases().items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **compute(pr))
        print("wrote", name)
