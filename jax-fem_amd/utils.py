"""Solution output and mesh input in VTK's XML unstructured-grid format (.vtu).

Replaces the reference's meshio-based ``save_sol`` / ``modify_vtu_file``
(jax_fem/utils.py:13-80) with a self-contained writer and reader:

* ``save_sol(fe, sol, sol_file, cell_infos, point_infos)`` writes the mesh, the
  nodal solution as point data ``sol`` (float32, as the reference) and optional
  named cell / point arrays, ASCII;
* ``read_vtu(path)`` returns a ``Mesh`` with ``points``, ``cells_dict`` keyed by
  meshio's cell-type names (e.g. 'hexahedron', or 'VTK_LAGRANGE_HEXAHEDRON' for
  FEniCSx output), ``point_data`` and ``cell_data``; ASCII and base64 ("binary",
  with or without zlib compression) data arrays;
* ``modify_vtu_file`` rewrites the version string 2.2 to 1.0 as the reference does.
"""
import base64
import os
import re
import struct
import zlib

import numpy as np

from .mesh import Mesh, get_meshio_cell_type

# meshio cell-type name <-> VTK cell type id
VTK_TYPES = {"vertex": 1, "line": 3, "triangle": 5, "quad": 9, "tetra": 10, "hexahedron": 12, "wedge": 13,
             "triangle6": 22, "quad8": 23, "tetra10": 24, "hexahedron20": 25, "quad9": 28, "hexahedron27": 29,
             "VTK_LAGRANGE_QUADRILATERAL": 70, "VTK_LAGRANGE_HEXAHEDRON": 72}
VTK_NAMES = {v: k for k, v in VTK_TYPES.items()}
_DTYPES = {"Float32": np.float32, "Float64": np.float64, "Int8": np.int8, "UInt8": np.uint8, "Int16": np.int16,
           "UInt16": np.uint16, "Int32": np.int32, "UInt32": np.uint32, "Int64": np.int64, "UInt64": np.uint64}
_VTK_NAME = {np.dtype(v): k for k, v in _DTYPES.items()}


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _array_xml(name, data, ncomp=None):
    data = np.ascontiguousarray(data)
    if ncomp is None:
        ncomp = 1 if data.ndim == 1 else int(np.prod(data.shape[1:]))
    body = " ".join(repr(float(x)) if data.dtype.kind == "f" else str(int(x)) for x in data.reshape(-1))
    attrs = 'type="%s"' % _VTK_NAME[data.dtype]
    if name:
        attrs += ' Name="%s"' % name
    if ncomp > 1:
        attrs += ' NumberOfComponents="%d"' % ncomp
    return '<DataArray %s format="ascii">%s</DataArray>\n' % (attrs, body)


def write_vtu(path, points, cells, cell_type, point_data=None, cell_data=None):
    """Write one cell block (connectivity (ncell, nn), meshio name cell_type) with data."""
    points = _host(points).astype(np.float64)
    if points.shape[1] == 2:
        points = np.hstack([points, np.zeros((len(points), 1))])
    cells = _host(cells).astype(np.int64)
    ncell, nn = cells.shape
    out = ['<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">\n',
           '<UnstructuredGrid>\n<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(points), ncell),
           "<Points>\n", _array_xml(None, points, 3), "</Points>\n<Cells>\n",
           _array_xml("connectivity", cells.reshape(-1)),
           _array_xml("offsets", (np.arange(1, ncell + 1) * nn).astype(np.int64)),
           _array_xml("types", np.full(ncell, VTK_TYPES[cell_type], dtype=np.uint8)), "</Cells>\n"]
    if point_data:
        out.append("<PointData>\n")
        for k, v in point_data.items():
            out.append(_array_xml(k, _host(v)))
        out.append("</PointData>\n")
    if cell_data:
        out.append("<CellData>\n")
        for k, v in cell_data.items():
            out.append(_array_xml(k, _host(v)))
        out.append("</CellData>\n")
    out.append("</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        f.write("".join(out))


def save_sol(fe, sol, sol_file, cell_infos=None, point_infos=None):
    """The reference's save_sol (utils.py:13-60): mesh + point data 'sol' (float32)."""
    sol = _host(sol)
    pd = {"sol": np.asarray(sol, dtype=np.float32)}
    cd = {}
    for name, data in (cell_infos or []):
        data = _host(data)
        assert data.shape == (fe.num_cells,), "cell data wrong shape, get %s, while num_cells = %d" % (
            data.shape, fe.num_cells)
        cd[name] = np.asarray(data, dtype=np.float32)
    for name, data in (point_infos or []):
        data = _host(data)
        assert len(data) == len(sol), "point data wrong shape!"
        pd[name] = np.asarray(data, dtype=np.float32)
    write_vtu(sol_file, fe.points, fe.cells, get_meshio_cell_type(fe.ele_type), pd, cd)


def modify_vtu_file(input_file_path, output_file_path):
    """Version 2.2 -> 1.0 in the VTKFile tag (utils.py:63-80)."""
    with open(input_file_path) as fin, open(output_file_path, "w") as fout:
        for line in fin:
            fout.write(line.replace('<VTKFile type="UnstructuredGrid" version="2.2">',
                                    '<VTKFile type="UnstructuredGrid" version="1.0">'))


def _decode(attrs, body, header_type, compressed):
    dt = np.dtype(_DTYPES[re.search(r'type="(\w+)"', attrs).group(1)]).newbyteorder("<")
    fmt = re.search(r'format="(\w+)"', attrs)
    fmt = fmt.group(1) if fmt else "ascii"
    if fmt == "ascii":
        return np.array(body.split(), dtype=np.float64).astype(dt)
    if fmt != "binary":
        raise NotImplementedError("VTU data format %r" % fmt)
    raw = base64.b64decode("".join(body.split()))
    hsz = 8 if header_type == "UInt64" else 4
    hf = "<Q" if hsz == 8 else "<I"
    if not compressed:
        n = struct.unpack(hf, raw[:hsz])[0]
        return np.frombuffer(raw[hsz:hsz + n], dtype=dt)
    nblocks = struct.unpack(hf, raw[:hsz])[0]
    sizes = struct.unpack("<%d%s" % (nblocks, hf[1]), raw[3 * hsz:3 * hsz + nblocks * hsz])
    pos = 3 * hsz + nblocks * hsz
    chunks = []
    for sz in sizes:
        chunks.append(zlib.decompress(raw[pos:pos + sz]))
        pos += sz
    return np.frombuffer(b"".join(chunks), dtype=dt)


def read_vtu(path):
    """A Mesh (points, cells of the first cell type, cells_dict, point_data, cell_data)."""
    text = open(path).read()
    head = re.search(r"<VTKFile([^>]*)>", text).group(1)
    ht = re.search(r'header_type="(\w+)"', head)
    header_type = ht.group(1) if ht else "UInt32"
    compressed = "compressor=" in head

    def arrays(section):
        m = re.search(r"<%s[^>]*>(.*?)</%s>" % (section, section), text, re.S)
        if not m:
            return {}
        out = {}
        for a in re.finditer(r"<DataArray([^>]*)>(.*?)</DataArray>", m.group(1), re.S):
            attrs, body = a.group(1), a.group(2)
            name = re.search(r'Name="([^"]*)"', attrs)
            nc = re.search(r'NumberOfComponents="(\d+)"', attrs)
            v = _decode(attrs, body, header_type, compressed)
            if nc and int(nc.group(1)) > 1:
                v = v.reshape(-1, int(nc.group(1)))
            out[name.group(1) if name else ""] = v
        return out
    pts = list(arrays("Points").values())[0].reshape(-1, 3).astype(np.float64)
    c = arrays("Cells")
    conn, offs, types = c["connectivity"].astype(np.int64), c["offsets"].astype(np.int64), c["types"].astype(int)
    starts = np.concatenate([[0], offs[:-1]])
    blocks = {}
    for t in np.unique(types):
        sel = np.flatnonzero(types == t)
        nn = int(offs[sel[0]] - starts[sel[0]])
        idx = starts[sel][:, None] + np.arange(nn)[None, :]
        blocks[VTK_NAMES.get(int(t), "vtk%d" % t)] = conn[idx]
    first = next(iter(blocks))
    mesh = Mesh(pts, blocks[first])
    mesh._cells_dict = blocks
    mesh.point_data = arrays("PointData")
    mesh.cell_data = arrays("CellData")
    return mesh
