"""On-disk formats of a particle set (the reference's Particles I/O,
src/Particles.h:553-684): the text stream of ``operator<<`` / ``operator>>``
and the VTK unstructured grid of ``copy_to_vtk_grid`` / ``copy_from_vtk_grid``
written as a VTK XML ``.vtu`` file.

Both work on host arrays (positions ``(n, D)``, ``id``, ``alive`` and named
per-particle columns of shape ``(n,)`` or ``(n, k)``); ``Particles.to_host()``
/ ``Particles.from_host()`` move a device particle set to and from them.

Text stream (src/Particles.h:553-571, src/Traits.h:433-455,
src/Vector.h:746-765): a header line naming the variables separated by
", ", then one line per particle with the values separated by ", " and
vectors written as "(x,y,z)".  ``precision=None`` prints like a default
``std::ostream`` (6 significant digits, ``%g``), as the reference does;
``precision=17`` round-trips FP64 exactly.  The reference's per-particle
random generator column has no counterpart here (its generators are not
part of this library) and is not written.

VTK (src/Particles.h:586-684): one VTK_VERTEX cell (type 1) per particle,
points padded to three components with zeros, and a point-data array per
variable other than position, as the reference's vtkFloatArray (Float32);
``float64=True`` keeps FP64 values (with Float32 arrays, as in the
reference, ids above 2^24 do not survive).  Points are written in Float64 (the
reference's vtkPoints default to FP32; positions here are FP64 throughout).
The reader takes files this writer produced (ascii or base64 binary,
uncompressed).
"""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np

__all__ = ["write_stream", "read_stream", "write_vtu", "read_vtu"]


def _fmt(v, precision):
    v = float(v)
    if precision is None:
        s = "%g" % v
    else:
        s = "%.*g" % (precision, v)
    return s


def _value_to_text(v, precision):
    v = np.asarray(v)
    if v.ndim == 0:
        if np.issubdtype(v.dtype, np.integer) or v.dtype == np.bool_:
            return str(int(v))
        return _fmt(v, precision)
    return "(" + ",".join(_value_to_text(e, precision) for e in v) + ")"


def write_stream(fh, position, id=None, alive=None, columns=None, precision=None):
    """Writes the ``operator<<`` layout to the text file object ``fh``."""
    position = np.asarray(position, dtype=np.float64)
    n = position.shape[0]
    id = np.arange(n, dtype=np.int64) if id is None else np.asarray(id)
    alive = np.ones(n, dtype=np.uint8) if alive is None else np.asarray(alive)
    columns = columns or {}
    names = ["position", "id", "alive"] + list(columns)
    fh.write(", ".join(names) + "\n")
    cols = [np.asarray(columns[k]) for k in columns]
    for i in range(n):
        vals = [_value_to_text(position[i], precision), str(int(id[i])), str(int(alive[i]))]
        vals += [_value_to_text(c[i], precision) for c in cols]
        fh.write(", ".join(vals) + "\n")


def _parse_value(tok):
    tok = tok.strip()
    if tok.startswith("("):
        return [float(t) for t in tok[1:-1].split(",")]
    return float(tok)


def read_stream(fh):
    """Reads a file written by ``write_stream``; returns (names, dict of
    arrays) with positions as float64 (n, D), id int64, alive uint8."""
    header = fh.readline().rstrip("\n")
    names = [h.strip() for h in header.split(", ")]
    rows = []
    for line in fh:
        line = line.rstrip("\n")
        if not line:
            continue
        # split on ", " outside parentheses
        toks, depth, cur = [], 0, ""
        i = 0
        while i < len(line):
            c = line[i]
            if c == "(":
                depth += 1
            elif c == ")":
                depth -= 1
            if depth == 0 and line.startswith(", ", i):
                toks.append(cur)
                cur = ""
                i += 2
                continue
            cur += c
            i += 1
        toks.append(cur)
        if len(toks) != len(names):
            raise ValueError("line has %d values for %d variables" % (len(toks), len(names)))
        rows.append([_parse_value(t) for t in toks])
    out = {}
    for k, name in enumerate(names):
        col = [r[k] for r in rows]
        if name == "position":
            out[name] = np.array(col, dtype=np.float64).reshape(len(rows), -1)
        elif name == "id":
            out[name] = np.array(col, dtype=np.int64)
        elif name == "alive":
            out[name] = np.array(col, dtype=np.uint8)
        else:
            out[name] = np.array(col, dtype=np.float64)
    return names, out


def _data_array(parent, name, arr, vtk_type, ncomp, binary):
    da = ET.SubElement(parent, "DataArray", type=vtk_type, Name=name,
                       NumberOfComponents=str(ncomp), format="binary" if binary else "ascii")
    flat = np.ascontiguousarray(arr).ravel()
    if binary:
        raw = flat.tobytes()
        da.text = base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")
    elif flat.dtype.kind == "f":
        da.text = " ".join(repr(float(v)) for v in flat)
    else:
        da.text = " ".join(str(int(v)) for v in flat)
    return da


def write_vtu(path, position, id=None, alive=None, columns=None, float64=False, binary=True):
    """Writes the particles as a VTK XML unstructured grid (``.vtu``)."""
    position = np.asarray(position, dtype=np.float64)
    n, dim = position.shape
    pts = np.zeros((n, 3), dtype="<f8")
    pts[:, :dim] = position
    id = np.arange(n, dtype=np.int64) if id is None else np.asarray(id)
    alive = np.ones(n, dtype=np.uint8) if alive is None else np.asarray(alive)
    ftype, fname = (np.dtype("<f8"), "Float64") if float64 else (np.dtype("<f4"), "Float32")
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0",
                      byte_order="LittleEndian", header_type="UInt64")
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(n), NumberOfCells=str(n))
    pdata = ET.SubElement(piece, "PointData")
    fields = [("id", id), ("alive", alive)] + list((columns or {}).items())
    for name, arr in fields:
        arr = np.asarray(arr)
        ncomp = 1 if arr.ndim == 1 else arr.shape[1]
        _data_array(pdata, name, arr.astype(ftype), fname, ncomp, binary)
    points = ET.SubElement(piece, "Points")
    _data_array(points, "Points", pts, "Float64", 3, binary)
    cells = ET.SubElement(piece, "Cells")
    _data_array(cells, "connectivity", np.arange(n, dtype="<i8"), "Int64", 1, binary)
    _data_array(cells, "offsets", np.arange(1, n + 1, dtype="<i8"), "Int64", 1, binary)
    _data_array(cells, "types", np.ones(n, dtype=np.uint8), "UInt8", 1, binary)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


_VTK_NP = {"Float32": "<f4", "Float64": "<f8", "Int64": "<i8", "Int32": "<i4",
           "UInt8": "u1", "UInt64": "<u8"}


def _read_array(da):
    dt = np.dtype(_VTK_NP[da.get("type")])
    ncomp = int(da.get("NumberOfComponents", "1"))
    text = (da.text or "").strip()
    if da.get("format") == "binary":
        raw = base64.b64decode(text)
        nbytes = struct.unpack("<Q", raw[:8])[0]
        arr = np.frombuffer(raw[8:8 + nbytes], dtype=dt)
    else:
        arr = np.array(text.split(), dtype=np.float64).astype(dt) if text else np.zeros(0, dt)
    return arr.reshape(-1, ncomp) if ncomp > 1 else arr


def read_vtu(path, dimension=3):
    """Reads a ``.vtu`` written by ``write_vtu``: returns a dict with
    ``position`` ((n, dimension) float64), ``id`` (int64), ``alive``
    (uint8) and every other point-data array."""
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    n = int(piece.get("NumberOfPoints"))
    pts = _read_array(piece.find("Points/DataArray")).reshape(n, 3)
    types = _read_array([d for d in piece.find("Cells") if d.get("Name") == "types"][0])
    if not np.all(types == 1):
        raise ValueError("not a particle grid: cells other than VTK_VERTEX")
    out = {"position": np.ascontiguousarray(pts[:, :dimension], dtype=np.float64)}
    for da in piece.find("PointData"):
        arr = _read_array(da)
        name = da.get("Name")
        if name == "id":
            arr = arr.astype(np.int64)
        elif name == "alive":
            arr = arr.astype(np.uint8)
        out[name] = arr
    return out
