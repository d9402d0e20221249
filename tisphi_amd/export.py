"""On-disk exports of the particle state.

CSV — the column set, order and header of the reference's export_csv
(eng/ui_sim.py:293-298, written with numpy.savetxt, comma-delimited).
VTU — the point cloud and every ``dump()`` data array in the layout of
pyevtk's pointsToVTK (eng/ui_sim.py:300-304): XML UnstructuredGrid header
and one raw appended-binary section.  pyevtk is not a dependency here, so
the file is written directly.
File stamps follow func_export_step / func_export_time (:306-323).
"""
import os

import numpy as np

CSV_COLUMNS = ["id0", "objId", "material", "pos.x", "pos.y", "pos.z", "vel.x", "vel.y", "vel.z",
               "density", "stress.xx", "stress.yy", "stress.zz", "stress.xy", "stress.yz",
               "stress.zx", "strain_equ"]


def step_stamp(count_step):
    return f"{count_step:06d}"


def time_stamp(cur_time):
    return f"time.secx1e6.{int(cur_time * 1e6):07d}"


def export_csv(path, pos, data):
    cols = []
    for c in CSV_COLUMNS:
        cols.append(pos[c] if c.startswith("pos.") else data[c])
    arr = np.array(cols).T
    np.savetxt(path, arr, delimiter=",", header=", ".join(CSV_COLUMNS))
    return path


_VTK_TYPE = {"float64": "Float64", "float32": "Float32", "int64": "Int64", "int32": "Int32",
             "uint8": "UInt8"}


def export_vtu(path_no_ext, pos, data):
    """Write ``path_no_ext + '.vtu'`` as pyevtk's pointsToVTK does
    (eng/ui_sim.py:300-304): an UnstructuredGrid of one VTK_VERTEX cell per
    particle, every array in one raw little-endian AppendedData section,
    each block preceded by its byte count as UInt64 (header_type UInt64),
    points interleaved x, y, z; Int32 connectivity / offsets, UInt8 types;
    point data in dump() order with the first array named as Scalars."""
    x = np.ascontiguousarray(pos["pos.x"], dtype=np.float64)
    y = np.ascontiguousarray(pos["pos.y"], dtype=np.float64)
    z = np.ascontiguousarray(pos["pos.z"], dtype=np.float64)
    n = x.size
    conn = np.arange(n, dtype="<i4")
    offs = np.arange(1, n + 1, dtype="<i4")
    types = np.ones(n, dtype=np.uint8)  # VTK_VERTEX
    arrays = [("points", np.stack([x, y, z], axis=1).astype("<f8"), 3)]
    cells = [("connectivity", conn, 1), ("offsets", offs, 1), ("types", types, 1)]
    pdata = [(k, np.ascontiguousarray(v).astype(np.asarray(v).dtype.newbyteorder("<")), 1)
             for k, v in data.items()]
    offset = 0

    def header(name, a, ncomp):
        nonlocal offset
        h = (f'\n<DataArray Name="{name}" NumberOfComponents="{ncomp}" '
             f'type="{_VTK_TYPE[a.dtype.name]}" format="appended" offset="{offset}"/>')
        offset += a.size * a.dtype.itemsize + 8
        return h

    xml = ['<?xml version="1.0"?>',
           '\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
           "\n<UnstructuredGrid>", f'\n<Piece NumberOfPoints="{n}" NumberOfCells="{n}">', "\n<Points>"]
    xml += [header(*arrays[0]), "\n</Points>", "\n<Cells>"]
    xml += [header(*c) for c in cells] + ["\n</Cells>"]
    if pdata:
        xml.append(f'\n<PointData Scalars="{pdata[0][0]}">')
        xml += [header(*d) for d in pdata] + ["\n</PointData>"]
    xml += ["\n</Piece>", "\n</UnstructuredGrid>", '\n<AppendedData encoding="raw">\n_']
    path = path_no_ext + ".vtu"
    with open(path, "wb") as f:
        f.write("".join(xml).encode())
        for _, a, _ in arrays + cells + pdata:
            f.write(np.uint64(a.size * a.dtype.itemsize).astype("<u8").tobytes())
            f.write(a.tobytes())
        f.write(b"\n</AppendedData>\n</VTKFile>")
    return path


def read_vtu_points(path):
    """Points and point-data arrays of a .vtu written by export_vtu (raw
    appended data, UInt64 block headers)."""
    import re
    raw = open(path, "rb").read()
    k = raw.index(b'<AppendedData encoding="raw">')
    head = raw[:k].decode()
    base = raw.index(b"_", k) + 1
    types = {"Float64": "<f8", "Float32": "<f4", "Int64": "<i8", "Int32": "<i4", "UInt8": "u1"}

    def block(off, dt):
        nb = int(np.frombuffer(raw, "<u8", 1, base + off)[0])
        return np.frombuffer(raw, types[dt], nb // np.dtype(types[dt]).itemsize, base + off + 8).copy()

    arrs = re.findall(r'<DataArray Name="([^"]+)" NumberOfComponents="(\d+)" type="(\w+)" '
                      r'format="appended" offset="(\d+)"/>', head)
    out = {name: (block(int(off), dt), int(nc)) for name, nc, dt, off in arrs}
    pts = out.pop("points")[0].reshape(-1, 3)
    pd_start = head.find("<PointData")
    names = re.findall(r'<DataArray Name="([^"]+)"', head[pd_start:]) if pd_start >= 0 else []
    return pts, {nm: out[nm][0] for nm in names}
