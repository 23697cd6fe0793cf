"""VTU files of CIPHER's output step (src/cipher.c:976-1091): at every
outputfreq-th step the reference writes "<outfile>_<step>.vtu" through
PETSc's VTK viewer with one value per output name per cell.

``write`` stores the leaves of a forest as an UnstructuredGrid of
axis-aligned pixels (2D) or voxels (3D) whose corners are shared points, and
each output name as a Float64 cell array of that name, base64-encoded with a
UInt64 byte-count header (VTK's "binary" inline format).  PETSc's writer
lays the same values out as one multi-component array; names, cell order and
point numbering are this module's own.  ``read`` parses such a file back.
"""
import base64
import xml.etree.ElementTree as ET

import numpy as np

VTK_PIXEL, VTK_VOXEL = 8, 11


def leaf_geometry(forest):
    """(points (npts, 3) float64, connectivity (ncells, 2^dim) int64, cell
    type): leaf corners in physical units, shared between leaves; corners in
    VTK pixel/voxel order (x fastest, then y, then z)."""
    dim = forest.dim
    offs = np.array([[(k >> d) & 1 for d in range(dim)] for k in range(1 << dim)], dtype=np.int64)
    ext = forest.cell_extent()
    corners = forest.anchor[:, None, :] + offs[None, :, :] * ext[:, None, None]
    flat = corners.reshape(-1, dim)
    uniq, inv = np.unique(flat, axis=0, return_inverse=True)
    dx = forest.size / forest.res
    pts = np.zeros((uniq.shape[0], 3), dtype=np.float64)
    pts[:, :dim] = uniq * dx[None, :]
    return pts, inv.reshape(forest.ncells, 1 << dim).astype(np.int64), (VTK_PIXEL if dim == 2 else VTK_VOXEL)


def _b64(a):
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def _array(parent, name, a, ncomp=1):
    types = {np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}
    e = ET.SubElement(parent, "DataArray", type=types[a.dtype], Name=name, format="binary")
    if ncomp > 1:
        e.set("NumberOfComponents", str(ncomp))
    e.text = _b64(a)


def write(path, forest, fields):
    """Write the leaves of forest and the cell fields {name: (ncells,)
    array} to path."""
    pts, conn, ctype = leaf_geometry(forest)
    n = forest.ncells
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0", byte_order="LittleEndian",
                      header_type="UInt64")
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(pts.shape[0]), NumberOfCells=str(n))
    ET.SubElement(piece, "Points")
    _array(piece.find("Points"), "Points", pts.ravel(), 3)
    cells = ET.SubElement(piece, "Cells")
    _array(cells, "connectivity", conn.ravel())
    _array(cells, "offsets", np.arange(1, n + 1, dtype=np.int64) * conn.shape[1])
    _array(cells, "types", np.full(n, ctype, dtype=np.uint8))
    cd = ET.SubElement(piece, "CellData")
    for name, v in fields.items():
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        if v.size != n:
            raise ValueError(f"field {name}: {v.size} values for {n} cells")
        _array(cd, name, v)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def read(path):
    """(points, connectivity, types, {name: cell array}) of a file written
    by ``write``."""
    root = ET.parse(path).getroot()
    dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}

    def arr(e):
        raw = base64.b64decode(e.text)
        nbytes = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
        return np.frombuffer(raw[8:8 + nbytes], dtype=dt[e.get("type")])

    piece = root.find("UnstructuredGrid/Piece")
    pts = arr(piece.find("Points/DataArray")).reshape(-1, 3)
    cells = {e.get("Name"): arr(e) for e in piece.find("Cells")}
    ncorner = int(cells["offsets"][0]) if cells["offsets"].size else 0
    conn = cells["connectivity"].reshape(-1, max(ncorner, 1))
    data = {e.get("Name"): arr(e) for e in piece.find("CellData")}
    return pts, conn, cells["types"], data
