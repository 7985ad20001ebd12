"""The file formats on either side of the path, through the C ABI (host code,
no GPU): mesh readers (.vtu, legacy .vtk), the VTU snapshot writer, TNL's
binary format for BoundaryConditionsStorage and vectors.

The reference holds no mesh, boundary or snapshot files (its meshes are LFS
pointers), so the checks are independent implementations of the published
formats written here with the standard library and numpy: VTK XML files in
every encoding a VTK writer produces (ascii, inline base64 with the header
encoded apart from or together with the data, zlib-compressed in several
blocks, raw and base64 appended data, big-endian), legacy VTK files in ASCII
and big-endian BINARY in both cell layouts, and the byte layout of TNL's
File/Array serialization (magic "TNLMN", int32-length type string, 64-bit
element count, raw elements), which is third-party code (TNL, not vendored)
and therefore pinned only by this restatement.
"""
import struct

import numpy as np
import pytest

from mhfem_amd import capi, meshgen, problems
from _formats import bc_bytes, legacy_vtk, mesh_arrays, pad3, parse_vtu, tnl_array, tnl_type, vtu_text


def tri_mesh():
    points, cells = meshgen.unit_square_triangles(3)
    return points, cells


def tet_mesh():
    points, cells = meshgen.unit_cube_kuhn(2)
    return points, cells


MESHES = {"tri": tri_mesh, "tet": tet_mesh, "quad": lambda: meshgen.unit_square_quads(3),
          "edge": lambda: meshgen.unit_interval(7), "hex": lambda: meshgen.unit_cube_hexes(2)}


VTU_CASES = {
    "ascii": dict(fmt="ascii"),
    "binary_u32_apart": dict(fmt="binary", header="UInt32"),
    "binary_u64_joint": dict(fmt="binary", header="UInt64", joint=True),
    "binary_u64_apart": dict(fmt="binary", header="UInt64"),
    "zlib_u32_blocks": dict(fmt="binary", header="UInt32", compress=True, blocksize=64),
    "zlib_u64_blocks": dict(fmt="binary", header="UInt64", compress=True, blocksize=48),
    "zlib_one_block": dict(fmt="binary", header="UInt64", compress=True),
    "appended_raw": dict(fmt="appended", header="UInt32"),
    "appended_raw_zlib": dict(fmt="appended", header="UInt64", compress=True, blocksize=40),
    "appended_base64": dict(fmt="appended", header="UInt32", appended_encoding="base64"),
    "appended_base64_zlib": dict(fmt="appended", header="UInt64", compress=True, blocksize=56,
                                 appended_encoding="base64"),
    "big_endian": dict(fmt="binary", header="UInt32", byte_order="BigEndian"),
    "big_endian_zlib": dict(fmt="binary", header="UInt64", byte_order="BigEndian", compress=True,
                            blocksize=64),
}


@pytest.mark.parametrize("mesh", sorted(MESHES))
@pytest.mark.parametrize("case", sorted(VTU_CASES))
def test_vtu_reader_every_encoding(tmp_path, mesh, case):
    points, cells = MESHES[mesh]()
    rng = np.random.default_rng(5)
    C = cells.shape[0]
    z0 = rng.standard_normal(C)
    z1 = rng.standard_normal(C).astype(np.float32)
    tag = rng.integers(-7, 100, C).astype(np.int32)
    p3, conn, offs, types = mesh_arrays(points, cells, np.int64 if "u64" in case else np.int32)
    data = {"InitialCondition[Z0]": ("Float64", z0), "InitialCondition[Z1]": ("Float32", z1),
            "tag": ("Int32", tag)}
    path = tmp_path / "m.vtu"
    path.write_bytes(vtu_text(p3, conn, offs, types, data, **VTU_CASES[case]))
    mf = capi.MeshFile(path)
    assert mf.dim == points.shape[1]
    np.testing.assert_array_equal(mf.points, points)
    np.testing.assert_array_equal(mf.cells, cells)
    assert mf.cell_data_names() == list(data)
    np.testing.assert_array_equal(mf.cell_data("InitialCondition[Z0]"), z0)
    np.testing.assert_array_equal(mf.cell_data("InitialCondition[Z1]"), z1.astype(np.float64))
    np.testing.assert_array_equal(mf.cell_data("tag"), tag.astype(np.float64))


def test_vtu_reader_float32_points_and_int64_types(tmp_path):
    points, cells = tet_mesh()
    p3, conn, offs, types = mesh_arrays(points, cells, np.int64)
    path = tmp_path / "m.vtu"
    path.write_bytes(vtu_text(p3.astype(np.float32), conn, offs, types.astype(np.int64), {},
                              fmt="binary", types_dtype="Int64"))
    mf = capi.MeshFile(path)
    np.testing.assert_array_equal(mf.points, points.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(mf.cells, cells)
    assert mf.cell_data_names() == []


def _expect_error(path, text, fmt="auto"):
    with pytest.raises(capi.MHFEMError) as e:
        capi.MeshFile(path, fmt)
    assert text in str(e.value), str(e.value)


def test_vtu_reader_rejects_bad_meshes(tmp_path):
    points, cells = tri_mesh()
    p3, conn, offs, types = mesh_arrays(points, cells)
    p = tmp_path / "a.vtu"
    # polygons (VTK_POLYGON) are not among the cell shapes of the scheme
    p.write_bytes(vtu_text(p3, conn[: 4 * 4], np.array([4, 8, 12, 16], np.int32), np.full(4, 7, np.uint8), {}))
    _expect_error(p, "unsupported cell type 7")
    t2 = types.copy()
    t2[3] = 10
    p.write_bytes(vtu_text(p3, conn, offs, t2, {}))
    _expect_error(p, "mixed cell types")
    c2 = conn.copy()
    c2[5] = p3.shape[0]
    p.write_bytes(vtu_text(p3, c2, offs, types, {}))
    _expect_error(p, "connectivity index out of range")
    q3 = p3.copy()
    q3[2, 2] = 0.5
    p.write_bytes(vtu_text(q3, conn, offs, types, {}))
    _expect_error(p, "z = 0 plane")
    good = vtu_text(p3, conn, offs, types, {})
    p.write_bytes(good.replace(b'type="UnstructuredGrid"', b'type="PolyData"', 1))
    _expect_error(p, "UnstructuredGrid only")
    # a corrupt base64 character inside the first inline array
    i = good.index(b'format="binary">\n') + len(b'format="binary">\n') + 10
    p.write_bytes(good[:i] + b"*" + good[i + 1:])
    _expect_error(p, "invalid base64")
    _expect_error(tmp_path / "missing.vtu", "cannot open file")
    _expect_error(tmp_path / "x.pvtu", "distributed meshes")
    _expect_error(tmp_path / "x.msh", "unsupported mesh format")
    p.write_bytes(good)
    mf = capi.MeshFile(p)
    with pytest.raises(KeyError):
        mf.cell_data("InitialCondition[Z0]")
    # the C entry point reports a missing array with a message
    L = capi.lib()
    h = L.mhfem_mesh_file_open(str(p).encode(), b"auto")
    try:
        assert L.mhfem_mesh_file_cell_data(h, b"nope", None, 0) == -7
        assert b"'nope' was not found" in L.mhfem_last_error(None)
    finally:
        L.mhfem_mesh_file_close(h)


@pytest.mark.parametrize("fmt", ["ascii", "binary", "zlib"])
@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_vtu_writer_follows_the_format(tmp_path, fmt, mesh):
    points, cells = MESHES[mesh]()
    rng = np.random.default_rng(11)
    C = cells.shape[0]
    z = rng.standard_normal(C) * 1e3
    w = np.arange(C, dtype=np.float64) / 7.0
    path = tmp_path / "s.vtu"
    capi.write_vtu(path, points, cells, {"Z": z, "w & <ratio>": w}, cycle=12, time=0.375, format=fmt)
    root, field, npts, nc, pts, cl, cdata = parse_vtu(path)
    assert root.get("byte_order") == "LittleEndian" and root.get("header_type") == "UInt64"
    assert (root.get("compressor") == "vtkZLibDataCompressor") == (fmt == "zlib")
    assert field["CYCLE"].tolist() == [12] and field["CYCLE"].dtype == np.int32
    assert field["TIME"].tolist() == [0.375]
    assert npts == points.shape[0] and nc == C
    np.testing.assert_array_equal(pts, pad3(points))
    nv = cells.shape[1]
    np.testing.assert_array_equal(cl["connectivity"], cells.ravel())
    np.testing.assert_array_equal(cl["offsets"], nv * np.arange(1, C + 1))
    np.testing.assert_array_equal(cl["types"], np.full(C, {(1, 2): 3, (2, 3): 5, (2, 4): 9, (3, 4): 10, (3, 8): 12}[
        (points.shape[1], nv)]))
    np.testing.assert_array_equal(cdata["Z"], z)  # %.17g text round-trips exactly
    np.testing.assert_array_equal(cdata["w & <ratio>"], w)
    # and the reader takes it back
    mf = capi.MeshFile(path)
    np.testing.assert_array_equal(mf.points, points)
    np.testing.assert_array_equal(mf.cells, cells)
    np.testing.assert_array_equal(mf.cell_data("Z"), z)


def test_vtu_writer_without_metadata_and_bad_input(tmp_path):
    points, cells = tri_mesh()
    path = tmp_path / "s.vtu"
    capi.write_vtu(path, points, cells)
    root, field, *_ = parse_vtu(path)
    assert field == {} and root.find("UnstructuredGrid/FieldData") is None
    bad = cells.copy()
    bad[0, 0] = points.shape[0]
    with pytest.raises(capi.MHFEMError, match="cell vertex out of range"):
        capi.write_vtu(path, points, bad)
    with pytest.raises(capi.MHFEMError, match="cannot open file"):
        capi.write_vtu(tmp_path / "no" / "dir.vtu", points, cells)
    with pytest.raises(ValueError):
        capi.write_vtu(path, points, cells, {"Z": np.zeros(3)})


@pytest.mark.parametrize("binary", [False, True])
@pytest.mark.parametrize("version", ["2.0", "5.1"])
@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_legacy_vtk_reader(tmp_path, binary, version, mesh):
    points, cells = MESHES[mesh]()
    rng = np.random.default_rng(3)
    C = cells.shape[0]
    z = rng.standard_normal(C)
    k = rng.random(C).astype(np.float32)
    path = tmp_path / "m.vtk"
    path.write_bytes(legacy_vtk(pad3(points), cells, {"InitialCondition[Z0]": z, "perm": k}, binary, version,
                                metadata=version.startswith("5")))
    mf = capi.MeshFile(path)
    assert mf.dim == points.shape[1]
    np.testing.assert_array_equal(mf.points, points)
    np.testing.assert_array_equal(mf.cells, cells)
    np.testing.assert_array_equal(mf.cell_data("InitialCondition[Z0]"), z)
    np.testing.assert_array_equal(mf.cell_data("perm"), k.astype(np.float64))
    assert sorted(mf.cell_data_names()) == ["InitialCondition[Z0]", "perm"]


def test_legacy_vtk_rejects_other_datasets(tmp_path):
    p = tmp_path / "g.vtk"
    p.write_bytes(b"# vtk DataFile Version 3.0\ngrid\nASCII\nDATASET STRUCTURED_POINTS\n")
    _expect_error(p, "UNSTRUCTURED_GRID only")
    p.write_bytes(b"not a vtk file\n")
    _expect_error(p, "is not a legacy VTK file")
    points, cells = tri_mesh()
    good = legacy_vtk(pad3(points), cells, {})
    p.write_bytes(good[:-40])
    with pytest.raises(capi.MHFEMError):
        capi.MeshFile(p)


def test_boundary_conditions_file_layout(tmp_path):
    rng = np.random.default_rng(2)
    n = 37
    bcs = capi.BoundaryConditionsStorage(n)
    bcs.tags[:] = rng.integers(0, 4, n)
    bcs.values[:] = rng.standard_normal(n)
    bcs.dirichletValues[:] = rng.standard_normal(n)
    path = tmp_path / "boundary.tnl"
    bcs.save(path)
    assert path.read_bytes() == bc_bytes(bcs.tags, bcs.values, bcs.dirichletValues)
    back = capi.BoundaryConditionsStorage.load(path)
    assert back.dofSize == n
    np.testing.assert_array_equal(back.tags, bcs.tags)
    np.testing.assert_array_equal(back.values, bcs.values)
    np.testing.assert_array_equal(back.dirichletValues, bcs.dirichletValues)


def test_boundary_conditions_file_like_init_py(tmp_path):
    # what examples/HeatEquation/init.py writes: every face FixedValue, zeros
    F = 12416
    path = tmp_path / "boundary.tnl"
    path.write_bytes(bc_bytes(np.zeros(F, np.uint8), np.zeros(F), np.zeros(F)))
    b = capi.BoundaryConditionsStorage.load(path)
    assert b.dofSize == F and not b.tags.any() and not b.values.any() and not b.dirichletValues.any()


def test_boundary_conditions_file_errors(tmp_path):
    p = tmp_path / "b.tnl"
    t, v, d = np.array([0, 1, 2, 3], np.uint8), np.ones(4), np.zeros(4)
    good = bc_bytes(t, v, d)
    p.write_bytes(good.replace(b"< double >", b"< float  >", 1))
    with pytest.raises(capi.MHFEMError, match="object type does not match"):
        capi.BoundaryConditionsStorage.load(p)
    p.write_bytes(b"XXXXX" + good[5:])
    with pytest.raises(capi.MHFEMError, match="wrong magic number"):
        capi.BoundaryConditionsStorage.load(p)
    # dofSize disagrees with the arrays
    p.write_bytes(tnl_type("TNL::MHFEM::BoundaryConditionsStorage< double >") + struct.pack("<q", 5)
                  + good[len(tnl_type("TNL::MHFEM::BoundaryConditionsStorage< double >")) + 8:])
    with pytest.raises(capi.MHFEMError, match="Invalid dofSize"):
        capi.BoundaryConditionsStorage.load(p)
    p.write_bytes(bc_bytes(np.array([0, 4], np.uint8), np.ones(2), np.ones(2)))
    with pytest.raises(capi.MHFEMError, match="invalid boundary condition type 4"):
        capi.BoundaryConditionsStorage.load(p)
    p.write_bytes(good[:-3])
    with pytest.raises(capi.MHFEMError, match="unexpected end of file"):
        capi.BoundaryConditionsStorage.load(p)
    bad = capi.BoundaryConditionsStorage(2)
    bad.tags[:] = [0, 9]
    with pytest.raises(capi.MHFEMError, match="invalid boundary condition type 9"):
        bad.save(p)


def test_boundary_conditions_older_array_type_strings(tmp_path):
    # arrays whose serialization type also names the device and the index type
    t, v, d = np.array([3, 0], np.uint8), np.array([1.5, 2.5]), np.array([-1.0, 4.0])

    def arr(elem, data):
        return tnl_type(f"TNL::Containers::Array< {elem}, [any_device], long >") + struct.pack("<q", len(data)) \
            + data.tobytes()

    p = tmp_path / "old.tnl"
    p.write_bytes(tnl_type("TNL::MHFEM::BoundaryConditionsStorage< double >") + struct.pack("<q", 2)
                  + arr("TNL::MHFEM::BoundaryConditionsType", t) + arr("double", v) + arr("double", d))
    b = capi.BoundaryConditionsStorage.load(p)
    np.testing.assert_array_equal(b.tags, t)
    np.testing.assert_array_equal(b.values, v)
    np.testing.assert_array_equal(b.dirichletValues, d)


def test_vector_file(tmp_path):
    x = np.random.default_rng(4).standard_normal(1001)
    p = tmp_path / "dof.vec.tnl"
    capi.save_vector(p, x)
    assert p.read_bytes() == tnl_array("double", x.astype("<f8"))
    np.testing.assert_array_equal(capi.load_vector(p), x)
    capi.save_vector(p, np.zeros(0))
    assert capi.load_vector(p).size == 0


# ---------------------------------------------------------------------------
# the reference's input files -> a problem (control.h::init, Solver::setInitialCondition)

def test_problem_from_files_matches_the_generated_problem(tmp_path):
    pr = problems.adr_2d(s=6, n=2)
    F, C = pr.num_faces, pr.num_cells
    rng = np.random.default_rng(9)
    z0 = rng.random(2 * C)
    mesh = tmp_path / "mesh.vtk"
    mesh.write_bytes(legacy_vtk(pad3(pr.points), pr.cells, {}, binary=True))
    ic = tmp_path / "initial.vtu"
    capi.write_vtu(ic, pr.points, pr.cells, {"InitialCondition[Z0]": z0[:C], "InitialCondition[Z1]": z0[C:]},
                   cycle=0)
    bcs = capi.BoundaryConditionsStorage(2 * F)
    bcs.tags[:], bcs.values[:], bcs.dirichletValues[:] = pr.tags, pr.bc_values, pr.bc_dirichlet
    bcs.save(tmp_path / "boundary.tnl")
    got = problems.from_files(str(mesh), str(tmp_path / "boundary.tnl"), str(ic), n=2, coeffs=pr.coeffs)
    np.testing.assert_array_equal(got.points, pr.points)
    np.testing.assert_array_equal(got.cells, pr.cells)
    np.testing.assert_array_equal(got.face_cells, pr.face_cells)
    np.testing.assert_array_equal(got.cell_faces, pr.cell_faces)
    np.testing.assert_array_equal(got.tags, pr.tags)
    np.testing.assert_array_equal(got.bc_values, pr.bc_values)
    np.testing.assert_array_equal(got.bc_dirichlet, pr.bc_dirichlet)
    np.testing.assert_array_equal(got.Z_iK0, z0)
    # defaults: zero Dirichlet everywhere and Z = 0 (Solver.hpp:281-294, :307-309)
    d = problems.from_files(str(mesh), n=2)
    assert (d.tags == 0).all() and not d.bc_dirichlet.any() and not d.Z_iK0.any() and d.tags.size == 2 * F
    with pytest.raises(ValueError, match="Wrong dofSize in BoundaryConditionsStorage"):
        problems.from_files(str(mesh), str(tmp_path / "boundary.tnl"), n=1)


def test_problem_from_files_with_quadrangles(tmp_path):
    pr = problems.adr_quads(s=5, n=2)
    mesh = tmp_path / "quads.vtu"
    capi.write_vtu(mesh, pr.points, pr.cells, format="binary")
    got = problems.from_files(str(mesh), n=2, coeffs=pr.coeffs)
    assert got.box and got.fpc == 4
    np.testing.assert_array_equal(got.cells, pr.cells)
    np.testing.assert_array_equal(got.cell_faces, pr.cell_faces)
    np.testing.assert_array_equal(got.face_cells, pr.face_cells)
