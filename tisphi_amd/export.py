"""On-disk exports of the particle state.

CSV — the column set, order and header of the reference's export_csv
(eng/ui_sim.py:293-298, written with numpy.savetxt, comma-delimited).
VTU — the point cloud and every ``dump()`` data array, as pyevtk's
pointsToVTK writes them (eng/ui_sim.py:300-304); pyevtk is not a dependency
here, so the XML UnstructuredGrid file is written directly (ASCII).
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


def export_vtu(path_no_ext, pos, data):
    """Write ``path_no_ext + '.vtu'``: one VTK_VERTEX cell per particle."""
    x, y, z = pos["pos.x"], pos["pos.y"], pos["pos.z"]
    n = len(x)
    path = path_no_ext + ".vtu"
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
                'header_type="UInt64">\n<UnstructuredGrid>\n')
        f.write(f'<Piece NumberOfPoints="{n}" NumberOfCells="{n}">\n')
        f.write('<Points>\n<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n')
        np.savetxt(f, np.stack([x, y, z], axis=1), fmt="%.17g")
        f.write('</DataArray>\n</Points>\n<Cells>\n')
        f.write('<DataArray type="Int64" Name="connectivity" format="ascii">\n')
        np.savetxt(f, np.arange(n, dtype=np.int64)[None], fmt="%d")
        f.write('</DataArray>\n<DataArray type="Int64" Name="offsets" format="ascii">\n')
        np.savetxt(f, np.arange(1, n + 1, dtype=np.int64)[None], fmt="%d")
        f.write('</DataArray>\n<DataArray type="UInt8" Name="types" format="ascii">\n')
        np.savetxt(f, np.ones((1, n), dtype=np.uint8), fmt="%d")
        f.write('</DataArray>\n</Cells>\n<PointData>\n')
        for k, v in data.items():
            v = np.asarray(v)
            t = "Int64" if np.issubdtype(v.dtype, np.integer) else "Float64"
            f.write(f'<DataArray type="{t}" Name="{k}" format="ascii">\n')
            np.savetxt(f, v[None], fmt="%d" if t == "Int64" else "%.17g")
            f.write('</DataArray>\n')
        f.write('</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n')
    return path


def read_vtu_points(path):
    """Points and point-data arrays of a file written by export_vtu."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    pts = np.array(piece.find("Points/DataArray").text.split(), dtype=np.float64).reshape(-1, 3)
    data = {}
    for da in piece.find("PointData"):
        dt = np.int64 if da.get("type") == "Int64" else np.float64
        data[da.get("Name")] = np.array(da.text.split(), dtype=dt)
    return pts, data
