"""Independent writers and parsers of the file formats around the path, for
the tests: VTK XML (.vtu) in every encoding a VTK writer produces, legacy VTK
(.vtk) in ASCII and big-endian BINARY, and TNL's binary object layout.  They
follow the published VTK file-format documentation and TNL's File/Array
serialization (magic "TNLMN", int32-length type string, 64-bit element count,
raw elements), written with the standard library and numpy only."""
import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

VTK_NP = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4", "UInt32": "u4",
          "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}


def pad3(points):
    p3 = np.zeros((points.shape[0], 3))
    p3[:, : points.shape[1]] = points
    return p3


# ---------------------------------------------------------------------------
# an independent VTK XML writer (VTK file-format documentation)

def _blocks(raw, hdr_t, compress, blocksize):
    """[header][payload] of one array in VTK's binary layout."""
    h = np.dtype(hdr_t)
    if not compress:
        return np.array([len(raw)], h).tobytes(), raw
    chunks = [raw[k:k + blocksize] for k in range(0, len(raw), blocksize)]
    z = [zlib.compress(c) for c in chunks]
    last = len(chunks[-1]) if len(chunks[-1]) < blocksize else 0
    header = np.array([len(chunks), blocksize, last] + [len(c) for c in z], h).tobytes()
    return header, b"".join(z)


def vtu_text(points3, conn, offsets, types, cell_data, fmt="binary", header="UInt32", compress=False,
             blocksize=1 << 15, joint=False, byte_order="LittleEndian", appended_encoding="raw",
             types_dtype=None):
    """A .vtu file.  fmt: ascii | binary | appended.  joint: header and data in
    one base64 stream (else encoded apart, as VTK does)."""
    end = "<" if byte_order == "LittleEndian" else ">"
    hdr_t = end + VTK_NP[header]
    appended = []

    def da(vtype, name, values, ncomp=1):
        arr = np.ascontiguousarray(values, np.dtype(end + VTK_NP[vtype]))
        attrs = f'type="{vtype}" Name="{name}"' + (f' NumberOfComponents="{ncomp}"' if ncomp > 1 else "")
        if fmt == "ascii":
            body = " ".join(repr(float(v)) if vtype.startswith("Float") else str(int(v)) for v in arr.ravel())
            return f'<DataArray {attrs} format="ascii">\n{body}\n</DataArray>\n'
        h, payload = _blocks(arr.tobytes(), hdr_t, compress, blocksize)
        if fmt == "binary":
            if joint:
                text = base64.b64encode(h + payload).decode()
            else:
                text = base64.b64encode(h).decode() + base64.b64encode(payload).decode()
            return f'<DataArray {attrs} format="binary">\n{text}\n</DataArray>\n'
        offset = sum(len(a) for a in appended)
        if appended_encoding == "raw":
            appended.append(h + payload)
        else:
            appended.append((base64.b64encode(h) + base64.b64encode(payload)))
        return f'<DataArray {attrs} format="appended" offset="{offset}"/>\n'

    out = ['<?xml version="1.0"?>\n<!-- written by the test -->\n',
           f'<VTKFile type="UnstructuredGrid" version="1.0" byte_order="{byte_order}" header_type="{header}"'
           + (' compressor="vtkZLibDataCompressor"' if compress else "") + ">\n<UnstructuredGrid>\n",
           f'<Piece NumberOfPoints="{points3.shape[0]}" NumberOfCells="{len(types)}">\n<Points>\n',
           da(points3.dtype.name.capitalize() if points3.dtype.kind == "f" else "Float64", "Points", points3, 3),
           "</Points>\n<Cells>\n",
           da({4: "Int32", 8: "Int64"}[conn.dtype.itemsize], "connectivity", conn),
           da({4: "Int32", 8: "Int64"}[offsets.dtype.itemsize], "offsets", offsets),
           da(types_dtype or "UInt8", "types", types),
           "</Cells>\n"]
    if cell_data:
        out.append("<CellData>\n")
        for name, (vtype, v) in cell_data.items():
            out.append(da(vtype, name, v))
        out.append("</CellData>\n")
    out.append("</Piece>\n</UnstructuredGrid>\n")
    blob = b"".join(o.encode() for o in out)
    if appended:
        blob += f'<AppendedData encoding="{appended_encoding}">\n_'.encode() + b"".join(appended) + \
            b"\n</AppendedData>\n"
    return blob + b"</VTKFile>\n"


def vtk_cell_type(dim, nv):
    """VTK_LINE 3, VTK_TRIANGLE 5, VTK_QUAD 9, VTK_TETRA 10, VTK_HEXAHEDRON 12."""
    return {(1, 2): 3, (2, 3): 5, (2, 4): 9, (3, 4): 10, (3, 8): 12}[(dim, nv)]


def mesh_arrays(points, cells, int_t=np.int32):
    nv = cells.shape[1]
    conn = cells.astype(int_t).ravel()
    offsets = (np.arange(1, cells.shape[0] + 1) * nv).astype(int_t)
    types = np.full(cells.shape[0], vtk_cell_type(points.shape[1], nv), np.uint8)
    return pad3(points), conn, offsets, types


# ---------------------------------------------------------------------------
# an independent parser of the library's VTU output

def parse_vtu(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "UnstructuredGrid"
    hdr = np.dtype("<" + VTK_NP[root.get("header_type", "UInt32")])
    compressed = root.get("compressor") == "vtkZLibDataCompressor"

    def decode(da):
        dt = np.dtype("<" + VTK_NP[da.get("type")])
        text = "".join(da.text.split())
        if da.get("format") == "ascii":
            return np.array(text and da.text.split(), dtype=np.float64).astype(dt)
        # header and data encoded apart: the header's base64 length is fixed
        if compressed:
            n0 = 4 * ((3 * hdr.itemsize + 2) // 3)
            nb = int(np.frombuffer(base64.b64decode(text[:n0]), hdr)[0])
            nh = 4 * (((3 + nb) * hdr.itemsize + 2) // 3)
            h = np.frombuffer(base64.b64decode(text[:nh]), hdr)
            z = base64.b64decode(text[nh:])
            out, at = b"", 0
            for cs in h[3:]:
                out += zlib.decompress(z[at:at + int(cs)])
                at += int(cs)
            assert len(out) == (int(h[0]) - 1) * int(h[1]) + (int(h[2]) or int(h[1]))
            return np.frombuffer(out, dt)
        nh = 4 * ((hdr.itemsize + 2) // 3)
        nbytes = int(np.frombuffer(base64.b64decode(text[:nh]), hdr)[0])
        raw = base64.b64decode(text[nh:])
        assert len(raw) == nbytes
        return np.frombuffer(raw, dt)

    grid = root.find("UnstructuredGrid")
    field = {da.get("Name"): decode(da) for da in grid.findall("FieldData/DataArray")}
    piece = grid.find("Piece")
    pts = decode(piece.find("Points/DataArray")).reshape(-1, 3)
    cells = {da.get("Name"): decode(da) for da in piece.findall("Cells/DataArray")}
    cdata = {da.get("Name"): decode(da) for da in piece.findall("CellData/DataArray")}
    return root, field, int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), pts, cells, cdata


# ---------------------------------------------------------------------------
# an independent legacy VTK writer

def legacy_vtk(points3, cells, cell_data, binary=False, version="2.0", extra_point_data=True,
               metadata=True):
    def block(values, dt):
        a = np.asarray(values)
        if binary:
            return np.ascontiguousarray(a, np.dtype(">" + dt)).tobytes() + b"\n"
        return (" ".join(str(v) if dt[0] in "iu" else repr(float(v)) for v in a.ravel()) + "\n").encode()

    C, nv = cells.shape
    out = [f"# vtk DataFile Version {version}\ntest mesh\n{'BINARY' if binary else 'ASCII'}\n"
           "DATASET UNSTRUCTURED_GRID\n".encode()]
    out += [f"POINTS {points3.shape[0]} double\n".encode(), block(points3, "f8")]
    if metadata:
        out.append(b"METADATA\nINFORMATION 0\n\n")
    if version.startswith("5"):
        out += [f"CELLS {C + 1} {C * nv}\nOFFSETS vtktypeint64\n".encode(),
                block(np.arange(C + 1) * nv, "i8"),
                b"CONNECTIVITY vtktypeint64\n", block(cells.ravel(), "i8")]
    else:
        lst = np.concatenate([np.full((C, 1), nv), cells], axis=1)
        out += [f"CELLS {C} {C * (nv + 1)}\n".encode(), block(lst, "i4")]
    dim = 1 if not points3[:, 1:].any() and nv == 2 else (2 if not points3[:, 2].any() else 3)
    out += [f"CELL_TYPES {C}\n".encode(), block(np.full(C, vtk_cell_type(dim, nv)), "i4")]
    if extra_point_data:
        out += [f"POINT_DATA {points3.shape[0]}\nSCALARS pid int 1\nLOOKUP_TABLE default\n".encode(),
                block(np.arange(points3.shape[0]), "i4")]
    out.append(f"CELL_DATA {C}\n".encode())
    scal = {k: v for k, v in cell_data.items() if k.startswith("InitialCondition")}
    for name, v in scal.items():
        out += [f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode(), block(v, "f8")]
    rest = {k: v for k, v in cell_data.items() if k not in scal}
    if rest:
        out.append(f"FIELD FieldData {len(rest)}\n".encode())
        for name, v in rest.items():
            out += [f"{name} 1 {C} float\n".encode(), block(v, "f4")]
    return b"".join(out)


# ---------------------------------------------------------------------------
# TNL's binary object format

def tnl_type(t):
    return b"TNLMN" + struct.pack("<i", len(t)) + t.encode()


def tnl_array(elem, data):
    return tnl_type(f"TNL::Containers::Array< {elem} >") + struct.pack("<Q", len(data)) + data.tobytes()


def bc_bytes(tags, values, dirichlet):
    return (tnl_type("TNL::MHFEM::BoundaryConditionsStorage< double >") + struct.pack("<q", tags.size)
            + tnl_array("TNL::MHFEM::BoundaryConditionsType", tags.astype(np.uint8))
            + tnl_array("double", values.astype("<f8")) + tnl_array("double", dirichlet.astype("<f8")))


