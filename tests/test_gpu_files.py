"""The reference's file-driven workflow on the MI355X path: input files as
examples/HeatEquation/init.py writes them (a mesh, "InitialCondition[Z0]" in a
.vtu, a BoundaryConditionsStorage .tnl), an INI file in the format of
examples/HeatEquation/config.ini, the C++ example run with it, and its VTU
snapshots checked against the oracle; and the linear-system dump of a failed
solve (Solver.hpp:1259-1264, :966-977)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io

from mhfem_amd import capi, meshgen, problems
from _formats import legacy_vtk, pad3, parse_vtu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "tnl-mhfem_amd", "examples", "heat_equation")


def init_py_initial_condition(points, cells, alpha=-0.05, beta=-0.05, gamma=5.0):
    """set_initial_condition of examples/HeatEquation/init.py:10-27 at the cell centres."""
    c = meshgen.cell_centroids(points, cells)
    x, y = c[:, 0] - 0.5, c[:, 1] - 0.5
    return np.maximum(0.0, (x * x / alpha + y * y / beta + gamma) * 0.2)


@pytest.mark.gpu
@pytest.mark.parametrize("solver,prec", [("gmres", "diagonal"), ("cg", "diagonal"), ("tfqmr", "ILU_ISAI"), ("idrs", "ilut"), ("sor", "none"),
                                         ("bicgstabl", "AMGX")])
def test_config_driven_heat_example_matches_oracle(tmp_path, solver, prec):
    import pyoracle

    points, cells = meshgen.unit_square_triangles(16)
    mesh = tmp_path / "mesh.vtk"
    mesh.write_bytes(legacy_vtk(pad3(points), cells, {}, binary=True))
    z0 = init_py_initial_condition(points, cells)
    capi.write_vtu(tmp_path / "initial.vtu", points, cells, {"InitialCondition[Z0]": z0}, cycle=0)
    # set_boundary_condition of init.py:30-44: every face FixedValue 0
    F = meshgen.face_numbering(cells, points.shape[0])[1].shape[0]
    capi.BoundaryConditionsStorage(F).save(tmp_path / "boundary.tnl")
    out = tmp_path / "output"
    out.mkdir()
    (tmp_path / "config.ini").write_text(f"""output-directory = {out}

# MHFEM space discretization
mesh = {mesh}
boundary-conditions-file = {tmp_path / 'boundary.tnl'}

# MHFEM time discretization
initial-condition = {tmp_path / 'initial.vtu'}
initial-time = 0
final-time = 0.3
snapshot-period = 0.1
time-step = 0.05

# MHFEM linear system solver
linear-solver = {solver}
gmres-restarting-max = 30
preconditioner = {prec}
convergence-residue = 1e-13

# Heat equation model parameters
diffusivity = 1e-3
""")
    print(f"case: heat example --config, s = 16, 6 steps, {solver} + {prec}", flush=True)
    res = subprocess.run([EXAMPLE, "--config", str(tmp_path / "config.ini")], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    assert "Total number of time steps:" in res.stdout

    pr = problems.from_files(str(mesh), str(tmp_path / "boundary.tnl"), str(tmp_path / "initial.vtu"), n=1,
                             coeffs=dict(N_ijK=np.array(1.0), u_ijKe=np.array(0.0), m_iK=np.array(1.0),
                                         D_ijK=np.array(1e-3), w_iKe=np.array(0.0), a_ijKe=np.array(0.0),
                                         r_ijK=np.array(0.0), f_iK=np.array(0.0)),
                             do_mobility_upwind=0)
    o = pyoracle.Oracle(pr)
    t = 0.0
    for step in range(4):
        if step:
            tau = min(0.1, 0.3 - t)
            o.run(t, t + tau, 0.05, tol=1e-13)
            t += tau
        snap = out / f"data_{step:05d}.vtu"
        root, field, npts, nc, pts, cl, cdata = parse_vtu(snap)
        assert root.get("compressor") == "vtkZLibDataCompressor"
        assert field["CYCLE"].tolist() == [step]
        assert abs(field["TIME"][0] - t) < 1e-12
        np.testing.assert_array_equal(pts[:, :2], points)
        np.testing.assert_array_equal(cl["connectivity"], cells.ravel())
        zo = o.array("Z_iK")
        err = np.linalg.norm(cdata["Z"] - zo) / np.linalg.norm(zo)
        print(f"snapshot {step}: rel err {err:.2e}", flush=True)
        assert err < 1e-10
        # the library's reader takes the snapshot back
        np.testing.assert_array_equal(capi.MeshFile(snap).cell_data("Z"), cdata["Z"])
    assert not (out / "data_00004.vtu").exists()


@pytest.mark.gpu
def test_failed_solve_saves_the_linear_system(tmp_path):
    pr = problems.adr_2d(s=8, n=2)
    print("case: adr2d s=8 n=2, 3 BiCGStab iterations to 1e-14", flush=True)
    s = capi.Solver(pr, tol=1e-14, max_iterations=3, output_directory=str(tmp_path))
    s.preIterate(0.0, pr.dt)
    s.assembleLinearSystem(0.0, pr.dt)
    with pytest.raises(capi.MHFEMError, match="did not converge"):
        s.solveLinearSystem()
    rowptr, col, val, rhs = s.export_system()
    A = scipy.io.mmread(str(tmp_path / "matrix.mtx")).tocsr()
    A.sort_indices()
    np.testing.assert_array_equal(A.indptr, rowptr)
    np.testing.assert_array_equal(A.indices, col)
    np.testing.assert_array_equal(A.data, val)  # %.17g round-trips every double
    np.testing.assert_array_equal(capi.load_vector(tmp_path / "rhs.vec.tnl"), rhs)
    np.testing.assert_array_equal(capi.load_vector(tmp_path / "dof.vec.tnl"), s.get_array("Z_iF"))
    s.close()
